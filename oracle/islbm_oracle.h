/*
 * islbm_oracle.h -- FP64 CPU restatement of the reference ISLBM stream-collide step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker / CPU baseline.
 * The product (the HIP library) never links or calls it.
 *
 * The restatement keeps the reference's padded single-rank layout
 * (NX6 = NX+7, NY6 = NY+7, NZ6 = NZ+6; index = J*NX6*NZ6 + K*NX6 + I,
 * C/variables.h:16-19, C/evolution.h:405) with jp = 1, so that the unique-node
 * layout and periodic indexing of the product are checked against an
 * independent addressing scheme.  Exchange with callers uses the unique-node
 * layout [q][k][j][i] (i = spanwise x fastest, k = wall-normal z slowest).
 *
 * Axis convention is the reference's: x = spanwise (LX), y = streamwise (LY),
 * z = wall-normal (LZ).
 */
#ifndef ISLBM_ORACLE_H
#define ISLBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_params {
    int    nx, ny, nz;      /* unique nodes: spanwise, streamwise, wall-normal */
    double lx, ly, lz;      /* domain (reference 4.5, 9.0, 3.036) */
    double cfl;             /* 0.6: minSize = (LZ-1)/NZ*CFL */
    double lxi;             /* 10.0: computational xi extent */
    double hill_alpha;      /* hill-length scale; 1 = reference profile */
    int    flat;            /* 1 = ChannelFunction (H = 0) */
    int    mesh_mode;       /* 0 = bisected a (reference), 1 = a = tanh(stretch/2) */
    double stretch;
    double tau;             /* > 0: tau given, Uref = Re*nu (reference form) */
    double re;
    double u_ref;           /* used when tau <= 0: nu = u_ref/Re, tau = 3nu/dt + 1/2 */
    double cs_smag;         /* Smagorinsky constant (0 = reference MRT) */
    int    ndtfrc;          /* forcing-controller cadence (steps) */
    int    xi_near_wall;    /* 0 = the reference's xi weights; 1 = departure xi clamped to the
                               node range and linear weights on the 4 levels next to each wall
                               (the library's opt-in ISLBM_XI_CLAMPED_LINEAR) */
} or_params;

typedef struct or_state or_state;

or_state* or_create(const or_params* p);
void      or_destroy(or_state* s);
int       or_threads(int n);                 /* set OpenMP threads, returns previous */

/* derived scalars: [dt, tau, nu, uref, a, dx, dy, force] */
void      or_scalars(const or_state* s, double* out8);

/* f = f_eq(rho0, u0 + U(-amp, amp) per node/component, SplitMix64(seed)) */
void      or_init_equilibrium(or_state* s, double rho0, const double u0[3], double amp, uint64_t seed);
void      or_set_f(or_state* s, const double* f_unique);          /* [19][nz][ny][nx] */
void      or_get_f(const or_state* s, double* f_unique);
void      or_get_macros(const or_state* s, double* rho, double* u, double* v, double* w);
void      or_step(or_state* s, int nsteps);
void      or_modify_forcing(or_state* s);
double    or_get_force(const or_state* s);
void      or_set_force(or_state* s, double f);
void      or_set_rho_modify(or_state* s, double dr);
double    or_flux_sum(const or_state* s);    /* current sum of the Ub accumulator */
long      or_step_count(const or_state* s);

/* Test entries that split the step (no reference counterpart):
 *   or_stream_only: interpolated streaming + bounce-back + BFL of every node, without the
 *                   collision, into the current populations;
 *   or_collide_one: the collision of one node (j, k) on given f_in[19] -> f_out[19],
 *                   mac[5] = [rho, u, v, w, tau_eff]. */
void      or_stream_only(or_state* s);
void      or_collide_one(const or_state* s, const double* f_in, int j, int k, double* f_out, double* mac);

/* Geometry tables, unique-node form (see islbm_oracle.c for layouts). */
int       or_get_table(const or_state* s, const char* name, double* out, int cap);

/* Stand-alone scalar routines restated from the reference (for golden-vector tests). */
double    or_hill(double y, double ly, double alpha, int flat);
double    or_lagrange6(double x, const double* nodes7, int m);
double    or_bisect_a(double lz, double minsize, int nz, double hill0);
double    or_delta_horizontal(double z_target, double y_large, double y_small, double y_point,
                              double ly, double alpha, double minsize);
double    or_delta_45(double z_pnt, double y_pnt, double y_zdom, double y_ydom,
                      double ly, double alpha, double minsize);
double    or_tanh_map(double L, double lattice, double a, int j, int n);

/* Turbulence statistics sums (statistics.h:122-224), unique-node read-out by name
 * (U V W P UU UV UW VV VW WW PU PV PW PP KT DUDX2 .. DWDZ2 UUU .. WWW). */
void      or_stats_enable(or_state* s);
void      or_stats_accumulate(or_state* s);
long      or_stats_count(const or_state* s);
int       or_stats_get(const or_state* s, const char* name, double* out_unique);
void      or_set_macros(or_state* s, const double* rho, const double* u, const double* v, const double* w);

/* D3Q19 MRT matrices as used by the oracle (row-major 19x19). */
void      or_mrt_matrices(double* M, double* Mi);

#ifdef __cplusplus
}
#endif
#endif
