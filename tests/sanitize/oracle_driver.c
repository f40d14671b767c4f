/* Drives the FP64 oracle (oracle/islbm_oracle.c) through its entry points on several
 * geometries for a sanitizer build: create, start, step, the split step, statistics,
 * every table, destroy.  Exit status 0 when every value it reads back is finite. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/islbm_oracle.h"

static int finite_all(const double* a, long n)
{
    for (long t = 0; t < n; ++t)
        if (!isfinite(a[t])) return 0;
    return 1;
}

int main(void)
{
    static const char* tables[] = {"x", "y", "xi", "z", "xw0", "xw2", "yw0", "yw2", "xiw3", "xiw4", "xiw5", "xiw6",
                                   "xiw15", "xiw16", "xiw17", "xiw18", "bfl_flag3", "bfl_q3", "bfl_flag16", "bfl_q16",
                                   "bfl_x37", "bfl_y378", "bfl_xi378", "bfl_xi16"};
    int bad = 0;
    for (int c = 0; c < 6; ++c) {
        or_params p = {8, 24 + 8 * c, 16 + 4 * c, 4.5, 9.0, 3.036, 0.6, 10.0, 1.0, 0, 0, 2.0, 0.8, 50.0, 0.0, 0.0, 5, 0};
        if (c == 1) p.flat = 1;
        if (c == 2) { p.mesh_mode = 1; p.xi_near_wall = 1; }
        if (c == 3) { p.hill_alpha = 0.5; p.ly = 5.142 + 3.858 * 0.5; p.cs_smag = 0.1; }
        if (c == 4) { p.hill_alpha = 3.0; p.ly = 5.142 + 3.858 * 3.0; p.xi_near_wall = 1; }
        if (c == 5) { p.tau = 0.0; p.u_ref = 0.05; p.re = 200.0; p.mesh_mode = 1; }
        or_state* s = or_create(&p);
        const double u0[3] = {0.0, 0.03, 0.0};
        or_init_equilibrium(s, 1.0, u0, 1e-3, 7 + c);
        or_step(s, 6);
        or_stream_only(s);
        const long n = (long)p.nx * p.ny * p.nz;
        double* f = malloc(sizeof(double) * 19 * n);
        double* m = malloc(sizeof(double) * 4 * n);
        or_get_f(s, f);
        or_get_macros(s, m, m + n, m + 2 * n, m + 3 * n);
        bad += !finite_all(f, 19 * n) + !finite_all(m, 4 * n);
        double fo[19], mac[5];
        or_collide_one(s, f, 1, 1, fo, mac);
        bad += !finite_all(fo, 19) + !finite_all(mac, 5);
        or_stats_enable(s);
        or_stats_accumulate(s);
        bad += or_stats_get(s, "DWDZ2", m) != 0 || !finite_all(m, n);
        const int cap = 7 * 2 * (p.ny + 8) * (p.nz + 8);
        double* t = malloc(sizeof(double) * cap);
        for (size_t k = 0; k < sizeof tables / sizeof *tables; ++k) {
            const int got = or_get_table(s, tables[k], t, cap);
            bad += got <= 0 || !finite_all(t, got);
        }
        or_modify_forcing(s);
        bad += !isfinite(or_get_force(s)) + !isfinite(or_flux_sum(s));
        free(t); free(f); free(m);
        or_destroy(s);
    }
    printf("oracle driver: %d problems\n", bad);
    return bad != 0;
}
