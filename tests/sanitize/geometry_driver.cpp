// Drives the library's host geometry code (csrc/islbm_geometry.cpp: mesh, Lagrange, BFL,
// flux and LES tables) and the host forms of the xi weights and the FP32 weight rounding
// (csrc/islbm_common.h) on several geometries, for a sanitizer build.  Exit status 0 when
// every table is finite and every FP32 weight set keeps its sum to within half an FP32 ulp of
// its largest weight (6e-8 of that weight's magnitude, at least 1).
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../lattie-boltzmann-method-note_of_paperreview_amd/csrc/islbm_internal.h"

int main()
{
    using namespace islbm;
    int bad = 0;
    for (int c = 0; c < 7; ++c) {
        islbm_params p{};
        p.nx = 8 + 4 * c; p.ny = 16 + 8 * c; p.nz = 12 + 6 * c;
        p.lx = 4.5; p.ly = 9.0; p.lz = 3.036; p.cfl = 0.6; p.lxi = 10.0; p.hill_alpha = 1.0;
        p.mesh_mode = c % 2; p.stretch = 2.0; p.tau = 0.7; p.re = 50.0; p.ndtfrc = 10;
        p.xi_near_wall = c % 3 == 0;
        if (c == 3) p.flat = 1;
        if (c == 4) { p.hill_alpha = 0.5; p.ly = 5.142 + 3.858 * 0.5; }
        if (c == 5) { p.hill_alpha = 3.0; p.ly = 5.142 + 3.858 * 3.0; }
        if (c == 6) { p.tau = 0.0; p.u_ref = 0.05; p.re = 5600.0; }
        HostGeom g;
        const std::string msg = build_host_geometry(p, g);
        if (!msg.empty()) {
            std::printf("case %d: %s\n", c, msg.c_str());
            ++bad;
            continue;
        }
        for (const std::vector<double>* v : {&g.x, &g.y, &g.xi, &g.zc, &g.xw0, &g.xw2, &g.yw0, &g.yw2, &g.bq, &g.bw,
                                             &g.fluxw, &g.flux_dx, &g.flux_dz, &g.les_delta})
            for (double d : *v) bad += !std::isfinite(d);
        const std::vector<double> zw = zderiv_weights(g);
        for (double d : zw) bad += !std::isfinite(d);
        // the per-node xi weights as the device computes them, and their FP32 sets
        for (int k = 0; k < p.nz; ++k)
            for (int j = 0; j < p.ny; ++j)
                for (int f = 0; f < NXI; ++f) {
                    const double pz = g.zc[(size_t)j * p.nz + k] + xi_dz(f) * g.minsize;
                    const double py = g.y[j] + xi_dy(f) * g.minsize;
                    const double H = hill_profile(py, p.ly, p.hill_alpha, p.flat);
                    const double pxi = p.lxi * (pz - (H + g.minsize / 2.0)) / (p.lz - H - g.minsize);
                    double w[7];
                    float w32[7];
                    xi_weights7(pxi, g.xi.data(), k, p.nz, p.xi_near_wall, w);
                    round_weights7(w, w32);
                    double s64 = 0.0, s32 = 0.0, big = 1.0;
                    for (int m = 0; m < 7; ++m) {
                        s64 += w[m];
                        s32 += (double)w32[m];
                        big = std::fmax(big, std::fabs(w[m]));
                    }
                    bad += !std::isfinite(s64) || std::fabs(s32 - s64) > 6e-8 * big;
                }
    }
    std::printf("geometry driver: %d problems\n", bad);
    return bad != 0;
}
