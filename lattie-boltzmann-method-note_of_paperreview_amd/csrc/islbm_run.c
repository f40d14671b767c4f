/* islbm_run -- the periodic-hill time loop of the reference's main program, in C, over the
 * library's C ABI (include/islbm.h).
 *
 * The reference's main.cu is not in its repository; its loop is reconstructed from the
 * notes (N/25...tex:653-694) and the headers it calls:
 *
 *   set-up   AllocateMemory, GenerateMesh_X/Y/Z, GetIntrplParameter_X/Y/Xi, BFLInitialization
 *            (memory.h:51-171, initialization.h:51-360)               -> islbm_create
 *   start    INIT == 0: InitialUsingDftFunc (initialization.h:6-49)    -> islbm_init_equilibrium
 *            INIT == 1: InitialUsingBkpData (fileIO.h:314-354)         -> islbm_checkpoint_read,
 *            with the step count it resumes from (--step0)             -> islbm_set_step_count
 *   loop     two steps per iteration, Launch_CollisionStreaming(ft, fd) then (fd, ft)
 *            (evolution.h:842); TBSWITCH: Launch_TurbulentSum after every step from step
 *            `tbstart` on (statistics.h:224); every NDTFRC steps Launch_ModifyForcingTerm
 *            (evolution.h:1003); every NDTMIT steps one Monitor.dat line (monitor.h:4-26);
 *            every `mass` steps the global mass correction that feeds rho_modify
 *            (evolution.h:597)
 *   end      OutputData / fileIO_PDF into <dir>/result (fileIO.h:52-67, 262-292) and, with
 *            statistics, Launch_OutputTB into <dir>/statistics (fileIO.h:379-451)
 *
 * Two ways to run the step:
 *   native   the library's own device state (unique-node layout), stepped by islbm_step,
 *            which also runs the forcing update every ndtfrc steps
 *   dropin   the reference's own calls on caller-owned padded FP64 device arrays:
 *            islbm_ref_init, Launch_CollisionStreaming(ft, fd) / (fd, ft), and
 *            Launch_ModifyForcingTerm every NDTFRC steps; the result files are the padded
 *            arrays as they stand, which is what the reference's OutputData writes
 *
 * The last output line is one JSON object: mode, grid, steps, seconds, MLUPS, the body
 * force and the last monitor values.  Exit status 0 on success, 1 on any failure (with the
 * library's error message on stderr).
 */
#define _POSIX_C_SOURCE 200809L /* clock_gettime */
#include <hip/hip_runtime_api.h>

#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../include/islbm.h"

#define NQ 19

/* D3Q19 velocities and weights in the reference's numbering (initialization.h:7-12) */
static const int kCx[NQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int kCy[NQ] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int kCz[NQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

static double weight(int q) { return q == 0 ? 1.0 / 3.0 : q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0; }

typedef struct {
    islbm_params p;
    int steps, ndtmit, tbstart, mass, init, dropin;
    long long step0;
    double u0[3], amp;
    uint64_t seed;
    const char* dir;
} options;

static int die(const char* what, const char* msg)
{
    fprintf(stderr, "islbm_run: %s: %s\n", what, msg ? msg : "");
    return 1;
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int parse_triple_i(const char* s, int32_t* a, int32_t* b, int32_t* c)
{
    return sscanf(s, "%d,%d,%d", a, b, c) == 3 ? 0 : -1;
}

static int parse_triple_d(const char* s, double v[3]) { return sscanf(s, "%lf,%lf,%lf", &v[0], &v[1], &v[2]) == 3 ? 0 : -1; }

static void usage(void)
{
    fprintf(stderr,
            "usage: islbm_run [--grid NX,NY,NZ] [--steps N] [--precision fp64|fp32|bf16|fp16]\n"
            "                 [--tau T] [--re R] [--uref U] [--cs C] [--alpha A] [--stretch S]\n"
            "                 [--ndtfrc N] [--ndtmit N] [--tbstart STEP] [--mass N] [--xi-clamped]\n"
            "                 [--u0 UX,UY,UZ] [--amp A] [--seed S] [--init 0|1] [--step0 N] [--dir DIR]\n"
            "                 [--dropin]\n");
}

static int parse(int argc, char** argv, options* o)
{
    islbm_default_params(&o->p);
    o->steps = 100;
    o->ndtmit = 0;
    o->tbstart = -1;
    o->mass = 0;
    o->init = 0;
    o->step0 = 0;
    o->dropin = 0;
    o->u0[0] = 0.0;
    o->u0[1] = 0.0;
    o->u0[2] = 0.0;
    o->amp = 0.0;
    o->seed = 20261203u;
    o->dir = "islbm_out";
    for (int a = 1; a < argc; ++a) {
        const char* k = argv[a];
        const char* v = a + 1 < argc ? argv[a + 1] : NULL;
        int used = 1;
        if (!strcmp(k, "--dropin")) {
            o->dropin = 1;
            used = 0;
        } else if (!strcmp(k, "--xi-clamped")) {
            o->p.xi_near_wall = ISLBM_XI_CLAMPED_LINEAR;
            used = 0;
        } else if (!v) {
            usage();
            return -1;
        } else if (!strcmp(k, "--grid")) {
            if (parse_triple_i(v, &o->p.nx, &o->p.ny, &o->p.nz)) return -1;
        } else if (!strcmp(k, "--steps")) {
            o->steps = atoi(v);
        } else if (!strcmp(k, "--precision")) {
            o->p.precision = !strcmp(v, "fp64") ? ISLBM_FP64 : !strcmp(v, "fp32") ? ISLBM_FP32
                           : !strcmp(v, "bf16") ? ISLBM_BF16 : !strcmp(v, "fp16") ? ISLBM_FP16 : -1;
            if (o->p.precision < 0) return -1;
        } else if (!strcmp(k, "--tau")) {
            o->p.tau = atof(v);
        } else if (!strcmp(k, "--re")) {
            o->p.re = atof(v);
        } else if (!strcmp(k, "--uref")) {
            o->p.u_ref = atof(v);
            o->p.tau = 0.0; /* tau follows from Re and Uref */
        } else if (!strcmp(k, "--cs")) {
            o->p.cs_smag = atof(v);
        } else if (!strcmp(k, "--alpha")) {
            o->p.hill_alpha = atof(v);
            o->p.ly = 5.142 + 3.858 * o->p.hill_alpha;
        } else if (!strcmp(k, "--stretch")) {
            o->p.mesh_mode = ISLBM_MESH_STRETCH;
            o->p.stretch = atof(v);
        } else if (!strcmp(k, "--ndtfrc")) {
            o->p.ndtfrc = atoi(v);
        } else if (!strcmp(k, "--ndtmit")) {
            o->ndtmit = atoi(v);
        } else if (!strcmp(k, "--tbstart")) {
            o->tbstart = atoi(v);
        } else if (!strcmp(k, "--mass")) {
            o->mass = atoi(v);
        } else if (!strcmp(k, "--u0")) {
            if (parse_triple_d(v, o->u0)) return -1;
        } else if (!strcmp(k, "--amp")) {
            o->amp = atof(v);
        } else if (!strcmp(k, "--seed")) {
            o->seed = strtoull(v, NULL, 10);
        } else if (!strcmp(k, "--init")) {
            o->init = atoi(v);
        } else if (!strcmp(k, "--step0")) {
            o->step0 = atoll(v);
        } else if (!strcmp(k, "--dir")) {
            o->dir = v;
        } else {
            usage();
            return -1;
        }
        a += used;
    }
    /* two steps per loop iteration; the cadences count in those steps */
    if (o->steps < 0 || o->steps % 2 || o->step0 < 0 || o->p.ndtfrc <= 0 || o->p.ndtfrc % 2 || o->ndtmit < 0 || o->ndtmit % 2 ||
        o->mass < 0 || o->mass % 2 || (o->tbstart >= 0 && o->tbstart % 2)) {
        fprintf(stderr, "islbm_run: steps and the cadences must be even (two steps per iteration)\n");
        return -1;
    }
    if (o->dropin && (o->p.precision != ISLBM_FP64 || o->init != 0 || o->step0 != 0 || o->tbstart >= 0 || o->mass > 0 || o->amp != 0.0)) {
        fprintf(stderr, "islbm_run: --dropin runs the reference's FP64 calls from the equilibrium start only\n");
        return -1;
    }
    return 0;
}

static int make_dirs(const char* dir)
{
    char sub[1024];
    snprintf(sub, sizeof sub, "%s/result", dir);
    if ((mkdir(dir, 0755) && errno != EEXIST) || (mkdir(sub, 0755) && errno != EEXIST)) return -1;
    return 0;
}

static void print_summary(const options* o, const char* mode, double secs, double force, const double mon[4])
{
    const double n = (double)o->p.nx * o->p.ny * o->p.nz;
    printf("{\"mode\": \"%s\", \"grid\": [%d, %d, %d], \"precision\": %d, \"steps\": %d, \"seconds\": %.6f, "
           "\"mlups\": %.1f, \"force\": %.17g, \"monitor_v_u_w\": [%.17g, %.17g, %.17g]}\n",
           mode, o->p.nx, o->p.ny, o->p.nz, o->p.precision, o->steps, secs, secs > 0 ? n * o->steps / secs / 1e6 : 0.0,
           force, mon[0], mon[1], mon[2]);
    fflush(stdout);
}

/* ------------------------------------------------------------------------------------ */
/* native: the library's solver handle                                                   */
/* ------------------------------------------------------------------------------------ */
static int run_native(const options* o)
{
    islbm_solver* s = NULL;
    char monitor[1024];
    snprintf(monitor, sizeof monitor, "%s/Monitor.dat", o->dir);
    if (make_dirs(o->dir)) return die("output directory", o->dir);
    if (islbm_create(&o->p, 0, &s)) return die("islbm_create", islbm_last_error());
    int rc = 0;
    double mon[4] = {0, 0, 0, 0}, force = 0.0, mean = 0.0, t0 = 0.0, t1 = 0.0;
    if (o->init == 0) {
        if (islbm_init_equilibrium(s, 1.0, o->u0, o->amp, o->seed)) rc = die("islbm_init_equilibrium", islbm_last_error());
    } else if (islbm_checkpoint_read(s, o->dir, 0)) {
        rc = die("islbm_checkpoint_read", islbm_io_last_error());
    }
    /* the step count a restart resumes from (the result files carry none) */
    if (!rc && o->step0 > 0 && islbm_set_step_count(s, o->step0)) rc = die("islbm_set_step_count", islbm_last_error());
    if (!rc && islbm_sync(s)) rc = die("islbm_sync", islbm_last_error());
    t0 = now_s();
    for (int step = 0; !rc && step < o->steps; step += 2) {
        if (o->tbstart >= 0 && step == o->tbstart && islbm_stats_enable(s, 1)) {
            rc = die("islbm_stats_enable", islbm_last_error());
            break;
        }
        if (islbm_step(s, 2)) {
            rc = die("islbm_step", islbm_last_error());
            break;
        }
        if (o->ndtmit > 0 && (step + 2) % o->ndtmit == 0 && islbm_monitor_append(s, monitor)) rc = die("monitor", islbm_io_last_error());
        if (o->mass > 0 && (step + 2) % o->mass == 0 && islbm_mass_correct(s, 1.0, &mean)) rc = die("islbm_mass_correct", islbm_last_error());
    }
    if (!rc && islbm_sync(s)) rc = die("islbm_sync", islbm_last_error());
    t1 = now_s();
    if (!rc && islbm_checkpoint_write(s, o->dir, 0)) rc = die("islbm_checkpoint_write", islbm_io_last_error());
    if (!rc && o->tbstart >= 0 && o->tbstart < o->steps && islbm_stats_write(s, o->dir, 0))
        rc = die("islbm_stats_write", islbm_io_last_error());
    if (!rc && (islbm_monitor(s, mon) || islbm_get_force(s, &force))) rc = die("islbm_monitor", islbm_last_error());
    if (!rc) print_summary(o, "native", t1 - t0, force, mon);
    islbm_destroy(s);
    return rc;
}

/* ------------------------------------------------------------------------------------ */
/* dropin: the reference's entry points on caller-owned padded arrays                    */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    long long NX6, NY6, NZ6, npad;
} padded_dims;

static long long pidx(const padded_dims* d, long long I, long long J, long long K) { return J * d->NX6 * d->NZ6 + K * d->NX6 + I; }

/* The monitor node (NX6/2, NYD6/2, NZ6/2) of Launch_Monitor (monitor.h:4-26): v, u, w and
 * the body force, one line of Monitor.dat in the library's format. */
static int dropin_monitor(const padded_dims* d, double* const mac[4], const double* force_d, int step, const char* path,
                          double mon[4])
{
    const long long at = pidx(d, d->NX6 / 2, d->NY6 / 2, d->NZ6 / 2);
    const int order[3] = {1, 0, 2}; /* v, u, w from the padded u, v, w */
    for (int t = 0; t < 3; ++t)
        if (hipMemcpy(&mon[t], mac[order[t]] + at, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (hipMemcpy(&mon[3], force_d, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (!path) return 0;
    FILE* fp = fopen(path, "a");
    if (!fp) return -1;
    fprintf(fp, "%d\t %.15lf\t %.15lf\t %.15lf\t %.15lf\n", step, mon[0], mon[1], mon[2], mon[3]);
    fclose(fp);
    return 0;
}

static int dump(const char* dir, const char* name, const double* dev, long long n, double* host)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/result/%s_0.bin", dir, name);
    if (hipMemcpy(host, dev, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    FILE* fp = fopen(path, "wb");
    if (!fp) return -1;
    const size_t w = fwrite(host, sizeof(double), (size_t)n, fp);
    fclose(fp);
    return w == (size_t)n ? 0 : -1;
}

static int run_dropin(const options* o)
{
    const int nx = o->p.nx, ny = o->p.ny, nz = o->p.nz;
    padded_dims d = {nx + 7, ny + 7, nz + 6, 0};
    d.npad = d.NX6 * d.NY6 * d.NZ6;
    char monitor[1024];
    snprintf(monitor, sizeof monitor, "%s/Monitor.dat", o->dir);
    if (make_dirs(o->dir)) return die("output directory", o->dir);
    if (islbm_ref_init(&o->p)) return die("islbm_ref_init", islbm_last_error());

    int rc = 0;
    double *ft[NQ] = {0}, *fd[NQ] = {0};
    double *mac[4] = {0}, *force_d = NULL, *rho_mod_d = NULL; /* u, v, w, rho_d */
    double* host = (double*)malloc((size_t)d.npad * sizeof(double));
    double mon[4] = {0, 0, 0, 0}, force = 0.0;
    if (!host) rc = die("malloc", "host buffer");
    if (!rc && islbm_ref_device_arrays(&mac[0], &mac[1], &mac[2], &mac[3], &force_d, &rho_mod_d))
        rc = die("islbm_ref_device_arrays", islbm_last_error());
    /* AllocateMemory (memory.h:63-66): 19 arrays per buffer */
    for (int q = 0; !rc && q < NQ; ++q)
        if (hipMalloc((void**)&ft[q], (size_t)d.npad * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&fd[q], (size_t)d.npad * sizeof(double)) != hipSuccess)
            rc = die("hipMalloc", "population arrays");
    /* InitialUsingDftFunc (initialization.h:6-49) with the velocity u0: f = f_eq(1, u0) at
     * every node of the x/y-padded interior levels, f = w_q (rest) in the z ghost layers,
     * which the step never reads; SendDataToGPU copies it into both buffers
     * (communication.h:245-258).  rho_d, u, v, w likewise. */
    const double uu = o->u0[0] * o->u0[0] + o->u0[1] * o->u0[1] + o->u0[2] * o->u0[2];
    for (int q = 0; !rc && q < NQ; ++q) {
        const double eu = kCx[q] * o->u0[0] + kCy[q] * o->u0[1] + kCz[q] * o->u0[2];
        const double feq = weight(q) * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu);
        for (long long J = 0; J < d.NY6; ++J)
            for (long long K = 0; K < d.NZ6; ++K)
                for (long long I = 0; I < d.NX6; ++I) host[pidx(&d, I, J, K)] = (K >= 3 && K < d.NZ6 - 3) ? feq : weight(q);
        if (hipMemcpy(ft[q], host, (size_t)d.npad * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fd[q], host, (size_t)d.npad * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = die("hipMemcpy", "initial populations");
    }
    for (int a = 0; !rc && a < 4; ++a) {
        const double in = a == 3 ? 1.0 : o->u0[a], ghost = a == 3 ? 1.0 : 0.0;
        for (long long J = 0; J < d.NY6; ++J)
            for (long long K = 0; K < d.NZ6; ++K)
                for (long long I = 0; I < d.NX6; ++I) host[pidx(&d, I, J, K)] = (K >= 3 && K < d.NZ6 - 3) ? in : ghost;
        if (hipMemcpy(mac[a], host, (size_t)d.npad * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = die("hipMemcpy", "initial fields");
    }
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = die("hipDeviceSynchronize", "set-up");
    /* the reference-named calls return nothing: a failure shows as a new error message */
    char err0[512];
    snprintf(err0, sizeof err0, "%s", islbm_last_error());
    const double t0 = now_s();
    for (int step = 0; !rc && step < o->steps; step += 2) {
        Launch_CollisionStreaming(ft, fd);
        Launch_CollisionStreaming(fd, ft);
        if ((step + 2) % o->p.ndtfrc == 0) Launch_ModifyForcingTerm();
        if (strcmp(islbm_last_error(), err0)) {
            rc = die("Launch_CollisionStreaming", islbm_last_error());
            break;
        }
        if (o->ndtmit > 0 && (step + 2) % o->ndtmit == 0 && dropin_monitor(&d, mac, force_d, step + 2, monitor, mon))
            rc = die("monitor", "device read or Monitor.dat");
    }
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = die("hipDeviceSynchronize", "loop");
    const double t1 = now_s();
    /* OutputData / fileIO_PDF: the padded arrays as they stand (ft holds the last f_new) */
    const char* mn[4] = {"u", "v", "w", "rho"};
    for (int a = 0; !rc && a < 4; ++a)
        if (dump(o->dir, mn[a], mac[a], d.npad, host)) rc = die("result", mn[a]);
    for (int q = 0; !rc && q < NQ; ++q) {
        char nm[8];
        snprintf(nm, sizeof nm, "f%d", q);
        if (dump(o->dir, nm, ft[q], d.npad, host)) rc = die("result", nm);
    }
    if (!rc && dropin_monitor(&d, mac, force_d, o->steps, NULL, mon)) rc = die("monitor", "device read");
    if (!rc) {
        char path[1024];
        snprintf(path, sizeof path, "%s/result/0_force.dat", o->dir);
        FILE* fp = fopen(path, "w");
        force = mon[3];
        if (!fp) {
            rc = die("result", path);
        } else {
            fprintf(fp, "%.15lf", force);
            fclose(fp);
        }
    }
    if (!rc) print_summary(o, "dropin", t1 - t0, force, mon);
    for (int q = 0; q < NQ; ++q) {
        if (ft[q]) hipFree(ft[q]);
        if (fd[q]) hipFree(fd[q]);
    }
    free(host);
    islbm_ref_shutdown();
    return rc;
}

int main(int argc, char** argv)
{
    options o;
    if (parse(argc, argv, &o)) return 2;
    return o.dropin ? run_dropin(&o) : run_native(&o);
}
