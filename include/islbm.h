/*
 * islbm.h -- C ABI of the MI355X (gfx950) ISLBM periodic-hill solver.
 *
 * The solver advances the interpolation-supplemented D3Q19 MRT lattice Boltzmann step of
 * the reference (7-point Lagrange interpolated streaming on a body-fitted tanh mesh over
 * the ERCOFTAC hill, flat-wall bounce-back, BFL at the hill, body force driven by a
 * bulk-flux controller) on one GPU with one device-resident domain.
 *
 * Axis convention is the reference's: x = spanwise (LX, nx nodes), y = streamwise (LY,
 * ny nodes), z = wall-normal (LZ, nz nodes).  Host arrays exchanged through this ABI hold
 * unique nodes only, as [k][j][i] with i fastest; population arrays are [19][k][j][i].
 * The periodic duplicate planes and ghost layers of the reference's padded layout
 * (variables.h:16-19) are not stored: x and y periodicity is resolved by indexing.
 *
 * Every function returns ISLBM_OK (0) or a negative ISLBM_ERR_* code and never exits the
 * process (the reference's CHECK_CUDA calls exit(1), common.h:6-16); the message of the
 * last failure on the calling thread is returned by islbm_last_error().
 * One solver is driven from one host thread at a time; distinct solvers are independent.
 */
#ifndef ISLBM_H
#define ISLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISLBM_OK 0
#define ISLBM_ERR_ARG (-1)
#define ISLBM_ERR_HIP (-2)
#define ISLBM_ERR_BUDGET (-3)
#define ISLBM_ERR_STATE (-4)
#define ISLBM_ERR_UNSUPPORTED (-5)

/* population storage / arithmetic */
#define ISLBM_FP64 0 /* FP64 storage and arithmetic (the reference's precision)        */
#define ISLBM_FP32 1 /* FP32 storage and arithmetic                                    */
#define ISLBM_BF16 2 /* BF16 storage of (f_i - w_i) * s_i, FP32 arithmetic; the step   */
                     /* rounds its stores stochastically (dither: node, step count)   */
#define ISLBM_FP16 3 /* FP16 storage of (f_i - w_i) * s_i, FP32 arithmetic; stores as  */
                     /* BF16's: stochastic rounding                                   */

/* wall-normal mesh */
#define ISLBM_MESH_BISECT 0  /* tanh parameter a bisected so that the first cell == minSize
                                (GetNonuniParameter, initializationTool.h:16-39)         */
#define ISLBM_MESH_STRETCH 1 /* a = tanh(stretch / 2)                                   */

/* wall-normal (xi) interpolation next to the walls */
#define ISLBM_XI_REFERENCE 0      /* the reference's 7-point Lagrange weights at every level,
                                     extrapolating where a departure point lies between a
                                     wall and the first or last node (initialization.h:185-258) */
#define ISLBM_XI_CLAMPED_LINEAR 1 /* opt-in stabilisation, not in the reference: the departure
                                     xi is clamped to [xi of the first node, xi of the last
                                     node] and the four levels next to each wall (whose 7-point
                                     stencil is one-sided) interpolate linearly between the
                                     bracketing nodes; all other levels as the reference    */

/* step kernel */
#define ISLBM_KERNEL_AUTO 0      /* FP64: two-pass; else tiled where supported, else per-node */
#define ISLBM_KERNEL_PERNODE 1   /* one thread per node, direct neighbour reads          */
#define ISLBM_KERNEL_TILED 2     /* column-marching, LDS-staged, sum-factorised          */
#define ISLBM_KERNEL_TWOPASS 3   /* FP64 only: x/y filter pass into a 16-direction buffer  */
                                 /* (16 doubles per node, counted by islbm_required_bytes), */
                                 /* then the xi sums, then walls + collision in place;      */
                                 /* ISLBM_ERR_UNSUPPORTED for the other precisions          */
#define ISLBM_KERNEL_TILED_WIDE 4 /* the tiled kernel reading each population array through */
                                  /* its own 64-bit base (wide addressing): chosen by AUTO and */
                                  /* TILED for grids whose 19 arrays exceed 4 GiB; selectable */
                                  /* for any grid the tiled kernel takes                      */

/* Run-time form of the reference's compile-time parameter set (variables.h:5-102). */
typedef struct islbm_params {
    int32_t nx, ny, nz;    /* unique nodes: NX (spanwise), NY (streamwise), NZ (wall-normal) */
    double lx, ly, lz;     /* LX, LY, LZ (variables.h:6-8)                               */
    double cfl;            /* CFL; minSize = dt = (LZ-1)/NZ*CFL (variables.h:22-23,76)    */
    double lxi;            /* LXi, computational wall-normal extent (variables.h:30)      */
    double hill_alpha;     /* hill-length scale factor; 1 = reference hill (model.h)      */
    int32_t flat;          /* 1: flat channel (ChannelFunction, model.h:70-73)            */
    int32_t mesh_mode;     /* ISLBM_MESH_*                                               */
    double stretch;        /* tanh stretch for ISLBM_MESH_STRETCH                         */
    double tau;            /* > 0: relaxation time given, Uref = Re*nu (variables.h:95-99) */
    double re;             /* Re (variables.h:41)                                        */
    double u_ref;          /* used when tau <= 0: nu = u_ref/Re, tau = 3 nu/dt + 1/2      */
    double cs_smag;        /* Smagorinsky constant; 0 = the reference MRT                 */
    int32_t ndtfrc;        /* forcing-controller cadence in steps (variables.h:48)        */
    int32_t precision;     /* ISLBM_FP64 / FP32 / BF16 / FP16                             */
    int32_t kernel;        /* ISLBM_KERNEL_*                                              */
    int32_t device;        /* HIP device ordinal                                          */
    int32_t xi_near_wall;  /* ISLBM_XI_REFERENCE (default) or ISLBM_XI_CLAMPED_LINEAR        */
} islbm_params;

typedef struct islbm_solver islbm_solver;

/* The reference's shipped parameter set (variables.h): 32 x 128 x 64, tau 0.6833, Re 50,
 * CFL 0.6, LXi 10, NDTFRC 10000, bisected mesh; FP64, automatic kernel, device 0. */
int islbm_default_params(islbm_params* p);

/* Device bytes a solver with these parameters allocates. */
int islbm_required_bytes(const islbm_params* p, size_t* bytes);

/* Computes the mesh and interpolation/BFL tables in FP64 (replaces AllocateMemory,
 * GenerateMesh_X/Y/Z, GetIntrplParameter_X/Y/Xi and BFLInitialization,
 * memory.h:51-171 and initialization.h:51-360) and allocates the device state.
 * Fails with ISLBM_ERR_BUDGET when the state exceeds hbm_budget bytes (0 = no limit). */
int islbm_create(const islbm_params* p, size_t hbm_budget, islbm_solver** out);
int islbm_destroy(islbm_solver* s);

/* f = f_eq(rho0, u0 + delta), delta_c ~ U(-amp, amp) from SplitMix64(seed) at counter
 * ((j*nz + k)*nx + i)*3 + c; macroscopic fields set to (rho0, u0 + delta) and the body
 * force reset to 8 nu Uref / LZ^2 * 5 (InitialUsingDftFunc, initialization.h:6-49). */
int islbm_init_equilibrium(islbm_solver* s, double rho0, const double u0[3], double amp, uint64_t seed);

/* Host <-> device transfer of the current populations (FP64 host view, [19][nz][ny][nx])
 * and macroscopic fields ([nz][ny][nx]; any pointer may be NULL).  Replace
 * SendDataToGPU / SendDataToCPU (communication.h:245-273). */
int islbm_upload_f(islbm_solver* s, const double* f);
int islbm_download_f(islbm_solver* s, double* f);
int islbm_download_macros(islbm_solver* s, double* rho, double* u, double* v, double* w);

/* Advances nsteps steps; each step is one Launch_CollisionStreaming (evolution.h:842-1001):
 * bulk-flux accumulation, the fused stream-collide kernel, and every ndtfrc steps the
 * forcing update (Launch_ModifyForcingTerm, evolution.h:1003-1042) on the device.
 * Work is queued on the solver's stream; islbm_sync waits for it. */
int islbm_step(islbm_solver* s, int32_t nsteps);
int islbm_sync(islbm_solver* s);
/* islbm_step through a HIP graph: the first call with a chunk length records the chunk's
 * launches (flux probe + stream-collide per step) into a graph, later calls of the same
 * length replay it with one launch.  nsteps must be even (the population buffers
 * alternate); a chunk that contains a forcing update is stepped as islbm_step does.
 * No counterpart in the reference (its step is launched kernel by kernel). */
int islbm_step_graph(islbm_solver* s, int32_t nsteps);
/* The solver's HIP stream (hipStream_t) and the device bytes it holds. */
int islbm_stream(islbm_solver* s, void** stream);
int islbm_device_bytes(islbm_solver* s, size_t* bytes);

/* Launch_ModifyForcingTerm (evolution.h:1003-1042) now, on the device. */
int islbm_modify_forcing(islbm_solver* s);
int islbm_get_force(islbm_solver* s, double* force);
int islbm_set_force(islbm_solver* s, double force);
/* rho_modify added to F0 before the moments (evolution.h:597). */
int islbm_set_rho_modify(islbm_solver* s, double drho);
/* Current value of the bulk-flux accumulator (sum of v*dx*dz over the probe plane). */
int islbm_get_flux_sum(islbm_solver* s, double* sum);

/* Host -> device macroscopic fields ([nz][ny][nx] doubles; NULL skips a field): the state a
 * restart needs beside the populations (the flux probe reads v of the previous step). */
int islbm_upload_macros(islbm_solver* s, const double* rho, const double* u, const double* v, const double* w);
/* The parameters the solver was created with. */
int islbm_get_params(islbm_solver* s, islbm_params* p);

/* Global mass correction: sets rho_modify (added to F0 before the moments of the next step,
 * evolution.h:597) to rho0 - mean(rho) of the last step, with the mean from a deterministic
 * FP64 device reduction.  The reference's computation of rho_modify lives in its absent
 * main; this restores the mean density to rho0 in one step. */
int islbm_mass_correct(islbm_solver* s, double rho0, double* mean_rho);
/* islbm_mass_correct without any device allocation by the library: no counterpart in the
 * reference.  For callers that cannot allocate (graph-captured or tight serving loops,
 * harnesses that forbid hipMalloc).  The reduction's partial sums live in one static
 * device array of the library per GPU (513 doubles, held until the process ends); it fits
 * every grid, since the partial count does not grow with the grid.  Results are identical
 * to islbm_mass_correct.  Calls that overlap on one GPU (solvers on different streams or
 * threads) share that array and race: serialise them. */
int islbm_mass_correct_noalloc(islbm_solver* s, double rho0, double* mean_rho);
/* Launch_Monitor (monitor.h:4-26): out = [v, u, w, force] at the reference's monitor node
 * (NX6/2, NYD6/2, NZ6/2 of the padded single-rank layout). */
int islbm_monitor(islbm_solver* s, double out[4]);

/* Turbulence statistics (Launch_TurbulentSum, statistics.h:122-249): FP64 sums over samples of
 * U V W P UU UV UW VV VW WW PU PV PW PP KT, the nine squared velocity derivatives
 * DUDX2 .. DWDZ2 and UUU UUV UUW VVU VVV VVW WWU WWV WWW, [nz][ny][nx] each.
 * enable allocates and zeroes them (33 * 8 * nodes bytes beside islbm_required_bytes);
 * every > 0 takes a sample after every `every`-th step of islbm_step, every = 0 only on
 * islbm_stats_accumulate.  d/dx, d/dy: the reference's 4th-order central differences
 * (statistics.h:188-194); d/dz: 5-point derivative weights on the node's vertical mesh line
 * over the reference's window (statistics.h:196-199). */
int islbm_stats_enable(islbm_solver* s, int32_t every);
int islbm_stats_disable(islbm_solver* s);
int islbm_stats_accumulate(islbm_solver* s);
int islbm_stats_reset(islbm_solver* s);
int islbm_stats_count(islbm_solver* s, int64_t* count);
int islbm_stats_set_count(islbm_solver* s, int64_t count);
/* names[0..32] = the accumulator names, in storage order; returns 33. */
int islbm_stats_names(const char** names, int32_t cap);
int islbm_stats_download(islbm_solver* s, const char* name, double* out);
int islbm_stats_upload(islbm_solver* s, const char* name, const double* in);

/* The reference's on-disk formats (fileIO.h), for interoperability with its dumps.
 * Padded single-rank arrays of (NX+7)*(NY+7)*(NZ+6) doubles, index J*NX6*NZ6 + K*NX6 + I.
 * checkpoint_write: dir/result/{rho,u,v,w,f0..f18}_<rank>.bin and, for rank 0,
 *   dir/result/0_force.dat (OutputData, fileIO_PDF, fileIO.h:52-67, 262-292).
 * checkpoint_read: the same files back into the solver (InitialUsingBkpData, fileIO.h:314-354).
 * stats_write: dir/statistics/NAME/NAME_<rank>.bin over the interior with its duplicate
 *   planes, k outer, i inner, and dir/statistics/accu.dat (Launch_OutputTB, fileIO.h:379-451).
 * stats_read: InitialTBUsingBkpData (fileIO.h:453-508).
 * monitor_append: one "step v u w force" line of Monitor.dat (monitor.h:4-26). */
int islbm_checkpoint_write(islbm_solver* s, const char* dir, int32_t rank);
int islbm_checkpoint_read(islbm_solver* s, const char* dir, int32_t rank);
int islbm_stats_write(islbm_solver* s, const char* dir, int32_t rank);
int islbm_stats_read(islbm_solver* s, const char* dir, int32_t rank);
int islbm_monitor_append(islbm_solver* s, const char* path);
const char* islbm_io_last_error(void);

/* Derived scalars: [dt, tau, nu, uref, a, dx, dy, force0]. */
int islbm_scalars(islbm_solver* s, double out[8]);
/* Steps taken since creation / the last init. */
int islbm_step_count(islbm_solver* s, int64_t* steps);
/* Sets the step count (for a restart: the reference's checkpoint files carry none).  The
 * count places the forcing-controller updates and, with 16-bit storage, dithers the
 * stochastic rounding of the step's stores, so a restart given the count it stopped at
 * continues bit for bit.  No counterpart in the reference (its main keeps the count). */
int islbm_set_step_count(islbm_solver* s, int64_t steps);
/* Kernel in use (ISLBM_KERNEL_PERNODE, ISLBM_KERNEL_TILED, ISLBM_KERNEL_TILED_WIDE or
 * ISLBM_KERNEL_TWOPASS). */
int islbm_kernel_in_use(islbm_solver* s, int32_t* kernel);
/* Columns x rows of node columns one workgroup of the tiled kernel owns: 64 x 4, 32 x 8
 * (grids whose rows fill 32-column tiles better than 64-column ones) or 64 x 2 (rows that
 * are not whole 16-byte granules); 0 x 0 for the other kernels.  Fixed by the grid.
 * No counterpart in the reference (its step launches one thread per node in NT = 32 thread
 * rows, evolution.h:853-854 and 940, variables.h:102). */
int islbm_tile_shape(islbm_solver* s, int32_t* cols, int32_t* rows);

/* Geometry table in unique-node form, as FP64 (names and layouts as the oracle's:
 * x y xi z xw0 xw2 yw0 yw2 xiw3..xiw18 bfl_flag* bfl_q* bfl_x37 ... bfl_xi16).
 * Returns the count written, or a negative error (-count needed if cap is too small). */
int islbm_get_table(islbm_solver* s, const char* name, double* out, int32_t cap);

/* Timing of nsteps queued steps with HIP events on the solver's stream:
 * out[0] = ms for all nsteps, out[1] = summed ms of the stream-collide kernel launches
 * alone (events recorded around each launch on the stream it runs on). */
int islbm_time_steps(islbm_solver* s, int32_t nsteps, double out[2]);

const char* islbm_last_error(void);

/* ---------------------------------------------------------------------------------- */
/* Continuous-batching serving loop.  No counterpart in the reference, whose (absent)   */
/* main runs one solve per program; it serves callers that run many periodic-hill       */
/* solves of different grids, Reynolds numbers and hill steepnesses on one GPU.         */
/* A request is one solve (params, seeded equilibrium start, step count).  Requests are  */
/* admitted in submission order while the device-memory budget (islbm_required_bytes    */
/* of every admitted solve) and the slot count allow; an admitted solve stays resident,  */
/* runs on its own stream in chunks replayed from a HIP graph (islbm_step_graph, two    */
/* chunks in flight) and is retired when done, returning its memory to the budget.      */
/* One batch is driven from one host thread.                                             */
/* ---------------------------------------------------------------------------------- */
#define ISLBM_REQ_QUEUED 0
#define ISLBM_REQ_ACTIVE 1
#define ISLBM_REQ_DONE 2
#define ISLBM_REQ_FAILED 3
typedef struct islbm_batch islbm_batch;
/* hbm_budget 0 = 90% of the device's free memory; chunk = steps per graph replay (even). */
int islbm_batch_create(int32_t device, size_t hbm_budget, int32_t max_active, int32_t chunk, islbm_batch** out);
int islbm_batch_destroy(islbm_batch* b);
/* Queues a solve: f = f_eq(1, u0 + delta) as islbm_init_equilibrium, then nsteps steps.
 * keep_fields != 0 keeps rho, u, v, w for islbm_batch_result.  A solve larger than the
 * whole budget is marked ISLBM_REQ_FAILED at once. */
int islbm_batch_submit(islbm_batch* b, const islbm_params* p, const double u0[3], double amp, uint64_t seed,
                       int64_t nsteps, int32_t keep_fields, int64_t* id);
/* One pass of the loop: retire finished solves, admit queued ones that fit, queue the next
 * chunks of every active solve; with wait != 0 and nothing retired or admitted, first block
 * until a chunk completes.  counts = [queued, active, done, failed] (may be NULL). */
int islbm_batch_advance(islbm_batch* b, int32_t wait, int64_t counts[4]);
/* state = ISLBM_REQ_*; times = [submit, admit, done] on the islbm_batch_now clock (s). */
int islbm_batch_status(islbm_batch* b, int64_t id, int32_t* state, double times[3], int64_t* steps_done);
/* Fields of a finished request submitted with keep_fields (released by the call). */
int islbm_batch_result(islbm_batch* b, int64_t id, double* rho, double* u, double* v, double* w);
/* [nodes, mean rho, max |u|, mean v] of a finished request. */
int islbm_batch_summary(islbm_batch* b, int64_t id, double out[4]);
int islbm_batch_bytes_in_use(islbm_batch* b, size_t* bytes);
/* out = [grouped chunks launched, times the grouped chunk graph was recorded]. */
int islbm_batch_group_stats(islbm_batch* b, int64_t out[2]);
const char* islbm_batch_last_error(void);
double islbm_batch_now(void);

/* Bandwidth of a plain float4 copy kernel over `bytes` (read + write counted), averaged
 * over `iters` launches: the measured HBM reference beside the vendor peak. */
int islbm_copy_bandwidth(int32_t device, size_t bytes, int32_t iters, double* gbps);

/* Host-only consistency check of the collision operator: the opposite-pair evaluation
 * the kernels use against the row-by-row MRT form (MRT_Process.h:4-179) on seeded states.
 * out = max |difference| for [FP64, FP64+LES, FP32, FP32+LES].  Touches no GPU. */
int islbm_selftest_collision(double out[4]);

/* ---------------------------------------------------------------------------------- */
/* Reference-named entry points (evolution.h:842, :1003).  They drive a process-wide   */
/* default solver created by islbm_ref_init and take the reference's padded FP64        */
/* population layout: 19 device arrays of (NX+7) * (NY+7) * (NZ+6) doubles, index      */
/* J*(NX+7)*(NZ+6) + K*(NX+7) + I, unique node (i,j,k) at (I,J,K) = (i+3, j+3, k+3)     */
/* (variables.h:16-19 with jp = 1, evolution.h:405).  After each call the ghost and     */
/* periodic-duplicate layers of f_new and of the padded u, v, w, rho are filled as      */
/* periodicSW and the single-rank halo exchange leave them (evolution.h:732-777,        */
/* 980-1000).                                                                            */
/* ---------------------------------------------------------------------------------- */
int islbm_ref_init(const islbm_params* p);
int islbm_ref_shutdown(void);
/* Device pointers of the padded macroscopic arrays u, v, w, rho_d and of Force_d,
 * rho_modify_d (memory.h:56-60, 155-159). */
int islbm_ref_device_arrays(double** u, double** v, double** w, double** rho, double** force, double** rho_modify);
void Launch_CollisionStreaming(double* f_old[19], double* f_new[19]);
void Launch_ModifyForcingTerm(void);

#ifdef __cplusplus
}
#endif
#endif
