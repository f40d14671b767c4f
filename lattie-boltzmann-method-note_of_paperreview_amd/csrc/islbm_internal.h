// islbm_internal.h -- host-side geometry, device state and kernel launchers of the solver.
#pragma once

#include "../../include/islbm.h"
#include "islbm_common.h"

#include <string>
#include <vector>

namespace islbm {

// FP64 geometry of one solve in unique-node form (what AllocateMemory and the
// initialization.h routines produce for one rank, minus ghosts and duplicates).
struct HostGeom {
    int nx = 0, ny = 0, nz = 0;
    double minsize = 0, dt = 0, tau = 0, nu = 0, uref = 0, a = 0, dx = 0, dy = 0, force0 = 0;
    std::vector<double> x, y, xi;        // x[nx] y[ny] xi[nz]
    std::vector<double> zc;              // node z, [ny][nz]
    std::vector<double> xw0, xw2;        // [nx][7]
    std::vector<double> yw0, yw2;        // [ny][7]
    std::vector<int> bflag;              // [NBFL][2][ny]
    std::vector<double> bq;              // [NBFL][2][ny]
    std::vector<double> bw;              // [NBW][2][ny][7]
    std::vector<double> fluxw;           // [nz][nx]  dx_i * dz_k on the flux probe plane
    std::vector<double> flux_dx;         // [nx]      dx_i of the probe plane
    std::vector<double> flux_dz;         // [nz]      dz_k of the probe plane
    std::vector<double> les_delta;       // [nz][ny]  (dx dy dz_local)^(1/3)
    int flux_plane = 0;                  // unique j of the probe plane (y = LY - 3 dy)
};

// Validates p and fills g (everything except the per-node xi weights, which are computed
// on the device by launch_xi_weights).  Returns an error message or "".
std::string build_host_geometry(const islbm_params& p, HostGeom& g);

// d/dz weights of the 5-point Lagrange interpolant on the vertical mesh line of (j, k),
// nodes z[j][n .. n+4] with n = k - 2 clamped to [0, nz - 5], evaluated at z[j][k]:
// [ny][nz][5].
std::vector<double> zderiv_weights(const HostGeom& g);

// Device-resident tables in compute precision C.
template <typename C>
struct DevTables {
    const C* xw0;  // [nx][7]
    const C* xw2;
    const C* yw0;  // [ny][7]
    const C* yw2;
    const C* xiw;  // [nz][ny][NXI][7]
    const int* bflag;  // [NBFL][2][ny]
    const C* bq;       // [NBFL][2][ny]
    const C* bw;       // [NBW][2][ny][7]
    const C* les_delta;  // [nz][ny]
};

// Population storage codec: FP64/FP32 store f directly; 16-bit formats store
// (f_q - w_q) * scale_q.
enum class Store { F64, F32, BF16, FP16 };

struct LaunchGeom {
    int nx, ny, nz;
    long long n;   // nodes
    long long qs;  // population stride: nodes from the array of direction q to that of q + 1 (>= n)
};
// Population stride of a grid of n nodes stored in `bytes`-byte values: n, or for arrays of
// 4 MiB and more, n rounded up to an odd multiple of 512 KiB (strides that are even
// multiples of large powers of two, or just short of them, run the stream-collide step
// measurably slower).
long long pop_stride(long long n, int bytes);

// Kernel launchers (islbm_kernels.hip).
hipError_t launch_xi_weights(hipStream_t st, double* out, const double* zc, const double* y,
                             const double* xi, int nx, int ny, int nz, double ly, double lz,
                             double lxi, double alpha, int flat, double minsize, int near_wall);
hipError_t launch_convert_f64(hipStream_t st, const double* in, void* out, long long n, int to_fp32);

struct StepLaunch {
    Store store;
    bool fp64;            // compute precision
    int kernel;           // ISLBM_KERNEL_PERNODE / _TILED
    LaunchGeom g;
    const void* fo;
    void* fn;
    void* rho;
    void* u;
    void* v;
    void* w;
    const void* tables;   // DevTables<C>*, host copy passed by value
    const double* force;
    const double* rho_modify;
    double tau, dt, cs_smag;
    const float* scale;   // per-direction 16-bit scale (device), or null
    void* hbuf;           // two-pass FP64: horizontally filtered directions 3-18, [16][n]
    // bulk-flux probe fused into the tiled kernel (null flux_part: probed by launch_flux_probe)
    double* flux_part;    // [2][flux_nslot] per-step partial sums, by step parity
    const double* flux_dx;  // [nx] probe weights dx_i
    const double* flux_dz;  // [nz] probe weights dz_k
    double* flux_acc;     // the flux accumulator
    int flux_plane;       // unique j of the probe plane
    int flux_par;         // parity of this step: writes flux_part[par], adds flux_part[1 - par]
    int flux_nslot;
    bool wide_addr;       // tiled kernel: one resource per population array (forced, or for grids that need it)
    // 16-bit storage: the solve's step counter in two slots by buffer parity (flux_par).  A
    // launch reads slot flux_par, and one thread of it writes slot 1 - flux_par = that + 1, so
    // eager, graph-replayed and grouped steps all see the solve's own step count.  Null for
    // the other storage formats.
    double* step_ctr;
};
hipError_t launch_stream_collide(hipStream_t st, const StepLaunch& L);
hipError_t launch_stream_copy(hipStream_t st, const void* src, void* dst, long long bytes);
struct InitScale {
    float s[NQ];
};
// f = f_eq(rho0, u0 + seeded delta) in the population format and rho, u, v, w (the four
// consecutive arrays at mac) in the compute precision; scale = 16-bit scale per direction.
hipError_t launch_init_equilibrium(hipStream_t st, Store store, void* f, long long qs, void* mac, long long n, int nx,
                                   int ny, int nz, double rho0, const double u0[3], double amp, uint64_t seed,
                                   const float* scale);
bool tiled_supported(const LaunchGeom& g, Store store, bool fp64);
// The grid's 19 population arrays (or rho, u, v, w) exceed one buffer resource's 4 GiB: the
// tiled kernel then reads each array through a resource of its own (wide addressing).
bool tiled_needs_wide(const LaunchGeom& g, Store store);
// Grouped stepping (serving): one launch advances up to 8 FP32 solves whose grids take the
// tiled kernel with 16-byte staging (tiled_groupable) by one step each; solve e takes part
// while step < limit[e].  Results equal those of launch_stream_collide for each solve.
bool tiled_groupable(const LaunchGeom& g, Store store);
// A second stream and two events for a group of mixed tile widths: the two launches of a
// step then run concurrently (forked from st, joined back into st).
struct GroupFork {
    hipStream_t st2;
    hipEvent_t fork, join;
};
hipError_t launch_tiled_group(hipStream_t st, const StepLaunch* L, const int* limit, int n, int step,
                              const GroupFork* fork = nullptr);
constexpr int kGroupMaxSolves = 8;
// Host side of grouped stepping (islbm_api.hip), for the serving loop:
//   group_eligible: the solve can join a group (FP32, 16-byte-staged tiled kernel, no
//                   per-step statistics, automatic forcing updates);
//   group_limit:    steps it may take in a chunk of `want` without passing a controller update;
//   group_parity:   its current population buffer;
//   group_record:   enqueue one chunk on st (capturable: host state untouched);
//   group_commit:   advance the host state by the chunk.
bool group_eligible(islbm_solver* s);
int group_limit(islbm_solver* s, int64_t want);
int group_parity(islbm_solver* s);
int group_record(hipStream_t st, islbm_solver* const* ss, const int* limit, int n, const GroupFork* fork = nullptr);
void group_commit(islbm_solver* const* ss, const int* limit, int n);
// Slots of one step's probe partials written by the tiled kernel (z chunks x x tiles).
int tiled_flux_slots(const LaunchGeom& g, Store store);
// columns x rows of a tiled-kernel workgroup's tile for this grid and storage
void tiled_tile_shape(const LaunchGeom& g, Store store, int* cols, int* rows);
// Sum of the current probe-plane flux sum_k,i v dx_i dz_k into part[0] and zeros into
// part[1 .. nslot-1]: the partials the tiled kernel's next step adds, after v was set
// from outside the step (start, upload, restart).
hipError_t launch_flux_refresh(hipStream_t st, const void* v, bool fp64, const double* dx, const double* dz,
                               double* part, int nslot, int nx, int ny, int nz, int jplane);

// Turbulence statistics (islbm_stats.hip): NSTAT FP64 accumulators [a][k][j][i], names in
// kStatNames; zdw = 5-point wall-normal derivative weights [j][k][5] (zderiv_weights).
constexpr int NSTAT = 33;
extern const char* const kStatNames[NSTAT];
hipError_t launch_stats(hipStream_t st, bool fp64, const void* mac, double* acc, const double* zdw, double dx, double dy,
                        int nx, int ny, int nz);
// deterministic sum of n values (compute precision) into out[0]; part holds kSumBlocks partials
constexpr int kSumBlocks = 512;
hipError_t launch_sum(hipStream_t st, bool fp64, const void* x, long long n, double* part, double* out);
// Device address, on `device` (current on the calling thread), of the library's static
// reduction scratch: kSumScratch doubles (kSumBlocks partials + the sum).
constexpr int kSumScratch = kSumBlocks + 1;
constexpr int kMaxDevices = 64;
hipError_t sum_scratch(int device, double** out);
// Summary of a solve's macroscopic fields mac = rho, u, v, w ([4][n], compute precision),
// deterministic fixed-order FP64 reductions: out = {sum rho, sum v, max |u| (NaN if any
// component is NaN)}; part holds 3 kSumBlocks partials.
constexpr int kSummaryScratch = 3 * kSumBlocks + 3;
hipError_t launch_field_summary(hipStream_t st, bool fp64, const void* mac, long long n, double* part, double* out);
// The same for a solver (islbm_api.hip): launched on its stream into part (kSummaryScratch
// device doubles), then copied to the host; waits for the solver's stream.
int field_summary(islbm_solver* s, double* part, double out[3]);

hipError_t launch_flux_probe(hipStream_t st, const void* v, bool fp64, const double* fluxw, double* part,
                             double* acc, int nx, int ny, int nz, int jplane);
hipError_t launch_force_update(hipStream_t st, double* force, double* acc, double inv_area,
                               int ndtfrc, double beta, double uref);

// Reference padded layout (NX6 x NY6 x NZ6, variables.h:16-19 with jp = 1) <-> unique
// nodes, FP64, for the reference-named entry points.  narr arrays are moved per call.
struct PtrSet {
    double* p[NQ + 4];
};
hipError_t launch_padded_to_unique(hipStream_t st, const PtrSet& pad, const PtrSet& uni, int nx, int ny, int nz,
                                   int narr);
hipError_t launch_unique_to_padded(hipStream_t st, const PtrSet& uni, const PtrSet& pad, int nx, int ny, int nz,
                                   int narr);

}  // namespace islbm
