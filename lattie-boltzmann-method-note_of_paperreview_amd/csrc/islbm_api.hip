// islbm_api.hip -- C ABI of the ISLBM periodic-hill solver (declared in include/islbm.h).
//
// Replaces the reference's global-state set-up and step orchestration: AllocateMemory
// (memory.h:51-171), the initialization.h set-up routines, InitialUsingDftFunc
// (initialization.h:6-49), SendDataToGPU/CPU (communication.h:245-273),
// Launch_CollisionStreaming (evolution.h:842-1001) and Launch_ModifyForcingTerm
// (evolution.h:1003-1042).  All state of a solve lives in one islbm_solver and stays
// resident on its GPU; there is no halo exchange because the domain is not split.
#include "islbm_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace islbm;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return fail(ISLBM_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                        \
    } while (0)

size_t store_bytes(int precision) { return precision == ISLBM_FP64 ? 8 : (precision == ISLBM_FP32 ? 4 : 2); }
size_t compute_bytes(int precision) { return precision == ISLBM_FP64 ? 8 : 4; }

Store store_of(int precision)
{
    switch (precision) {
    case ISLBM_FP64: return Store::F64;
    case ISLBM_FP32: return Store::F32;
    case ISLBM_BF16: return Store::BF16;
    default: return Store::FP16;
    }
}

// 16-bit deviation scale per direction: (f - w_q) / w_q is O(Ma) for every q.
float dev_scale(int precision, int q) { return precision >= ISLBM_BF16 ? store_scale16(q) : 1.0f; }

uint16_t f32_to_bf16_rne(float v)
{
    uint32_t b;
    std::memcpy(&b, &v, 4);
    const uint32_t r = b + 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
float bf16_to_f32(uint16_t h)
{
    uint32_t b = (uint32_t)h << 16;
    float v;
    std::memcpy(&v, &b, 4);
    return v;
}

size_t table_bytes(const islbm_params& p)
{
    const size_t cb = compute_bytes(p.precision);
    const size_t nx = p.nx, ny = p.ny, nz = p.nz;
    size_t b = 0;
    b += 2 * nx * 7 * cb + 2 * ny * 7 * cb;          // x, y weights
    b += nz * ny * NXI * 7 * cb;                     // xi weights
    b += NBFL * 2 * ny * (sizeof(int) + cb);         // BFL flags, q
    b += (size_t)NBW * 2 * ny * 7 * cb;              // BFL weights
    b += nz * ny * cb;                               // LES filter width
    b += nz * nx * 8 + nz * 8;                       // flux weights, per-level flux partials
    b += 8 * 8;                                      // scalars
    b += (kSumBlocks + 1) * 8;                       // reduction partials
    return b;
}

}  // namespace

long long islbm::pop_stride(long long n, int bytes)
{
    constexpr long long H = 1LL << 19;  // 512 KiB
    const long long b = n * bytes;
    if (b < 8 * H) return n;
    long long m = (b + H - 1) / H;
    if (m % 2 == 0) ++m;
    return m * H / bytes;
}

namespace {

// the two-pass FP64 kernel: AUTO's choice for FP64
bool uses_twopass(const islbm_params& p)
{
    return p.precision == ISLBM_FP64 && (p.kernel == ISLBM_KERNEL_AUTO || p.kernel == ISLBM_KERNEL_TWOPASS);
}

size_t state_bytes(const islbm_params& p)
{
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    const size_t qs = (size_t)pop_stride((long long)n, (int)store_bytes(p.precision));
    return 2 * NQ * qs * store_bytes(p.precision) + 4 * n * compute_bytes(p.precision) + table_bytes(p) +
           (uses_twopass(p) ? 16 * qs * sizeof(double) : 0);  // H rows at the population stride
}

// FP32 form of consecutive 7-point weight sets (round_weights7)
__global__ void k_round_weights7(const double* in, float* out, long long ngroups)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ngroups) round_weights7(in + 7 * t, out + 7 * t);
}

}  // namespace

struct islbm_solver {
    islbm_params p{};
    HostGeom g;
    Store store = Store::F64;
    bool fp64 = true;
    int kernel = ISLBM_KERNEL_PERNODE;
    bool wide_addr = false;  // tiled kernel with one resource per population array
    hipStream_t stream = nullptr;
    long long n = 0;
    long long qs = 0;  // population stride (pop_stride)
    void* f[2] = {nullptr, nullptr};
    int cur = 0;
    void* mac[4] = {nullptr, nullptr, nullptr, nullptr};
    void* hbuf = nullptr;  // two-pass FP64: filtered directions 3-18, [16][n]
    // tables (compute precision)
    void *xw0 = nullptr, *xw2 = nullptr, *yw0 = nullptr, *yw2 = nullptr, *xiw = nullptr;
    void *bq = nullptr, *bw = nullptr, *les = nullptr;
    int* bflag = nullptr;
    double* fluxw = nullptr;
    double* scal = nullptr;  // [0] force, [1] rho_modify, [2] flux accumulator, [3..4] 16-bit step counters
    DevTables<double> td{};
    DevTables<float> tf{};
    float scale[NQ];
    long long steps = 0;
    bool auto_force = true;  // run the forcing controller every ndtfrc steps
    size_t bytes = 0;
    std::vector<hipEvent_t> evpool;
    // islbm_step_graph: the captured chunk per starting buffer parity (a graph's launches
    // name the population buffers of the parity it was recorded at)
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    // bulk-flux probe fused into the tiled kernel: per-step partials by step parity and the
    // probe weights dx [nx] followed by dz [nz]
    double* flux_part = nullptr;
    double* flux_dxz = nullptr;
    int flux_nslot = 0;
    double* red = nullptr;           // kSumBlocks partials + the sum (mass correction)
    // turbulence statistics (islbm_stats_enable)
    double* st_acc = nullptr;        // [NSTAT][n] FP64 sums
    double* st_zdw = nullptr;        // [ny][nz][5] wall-normal derivative weights
    int st_every = 0;
    long long st_count = 0;
};

namespace {

void free_all(islbm_solver* s)
{
    void* ptrs[] = {s->f[0], s->f[1], s->mac[0], s->hbuf, s->xw0, s->xw2, s->yw0,
                    s->yw2, s->xiw, s->bq, s->bw, s->les, s->bflag, s->fluxw, s->scal};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    for (hipEvent_t e : s->evpool) (void)hipEventDestroy(e);
    for (hipGraphExec_t g : s->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (void* q : {(void*)s->red, (void*)s->st_acc, (void*)s->st_zdw, (void*)s->flux_part, (void*)s->flux_dxz})
        if (q) (void)hipFree(q);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

template <typename C>
int upload_table(void** dst, const std::vector<double>& src, bool weights7 = false)
{
    std::vector<C> tmp(src.begin(), src.end());
    if constexpr (sizeof(C) == 4)
        if (weights7)  // consecutive 7-point weight sets: rounded to keep their sums
            for (size_t g = 0; g + 7 <= src.size(); g += 7) round_weights7(&src[g], &tmp[g]);
    HIPCHK(hipMalloc(dst, tmp.size() * sizeof(C) + 8));
    HIPCHK(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(C), hipMemcpyHostToDevice));
    return ISLBM_OK;
}

template <typename C>
int upload_tables(islbm_solver* s)
{
    const HostGeom& g = s->g;
    int rc;
    if ((rc = upload_table<C>(&s->xw0, g.xw0, true)) || (rc = upload_table<C>(&s->xw2, g.xw2, true)) ||
        (rc = upload_table<C>(&s->yw0, g.yw0, true)) || (rc = upload_table<C>(&s->yw2, g.yw2, true)) ||
        (rc = upload_table<C>(&s->bq, g.bq)) || (rc = upload_table<C>(&s->bw, g.bw, true)) ||
        (rc = upload_table<C>(&s->les, g.les_delta)))
        return rc;
    // BFL branch codes decided in FP64 (see apply_walls)
    std::vector<int> code(g.bflag.size());
    for (size_t t = 0; t < code.size(); ++t)
        code[t] = !g.bflag[t] ? 0 : (g.bq[t] > 0.5 ? 1 : (g.bq[t] < 0.5 ? 2 : 3));
    HIPCHK(hipMalloc((void**)&s->bflag, code.size() * sizeof(int)));
    HIPCHK(hipMemcpy(s->bflag, code.data(), code.size() * sizeof(int), hipMemcpyHostToDevice));
    // flux weights [nz][nx], followed by nz per-level partial sums of the flux probe
    HIPCHK(hipMalloc((void**)&s->fluxw, (g.fluxw.size() + g.nz) * sizeof(double)));
    HIPCHK(hipMemcpy(s->fluxw, g.fluxw.data(), g.fluxw.size() * sizeof(double), hipMemcpyHostToDevice));

    // per-node xi weights: FP64 on the device, then stored in compute precision
    const size_t nxi = (size_t)g.nz * g.ny * NXI * 7;
    double *d64 = nullptr, *dzc = nullptr, *dy = nullptr, *dxi = nullptr;
    HIPCHK(hipMalloc((void**)&d64, nxi * sizeof(double)));
    HIPCHK(hipMalloc((void**)&dzc, g.zc.size() * sizeof(double)));
    HIPCHK(hipMalloc((void**)&dy, g.y.size() * sizeof(double)));
    HIPCHK(hipMalloc((void**)&dxi, g.xi.size() * sizeof(double)));
    HIPCHK(hipMemcpy(dzc, g.zc.data(), g.zc.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dy, g.y.data(), g.y.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dxi, g.xi.data(), g.xi.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(launch_xi_weights(s->stream, d64, dzc, dy, dxi, g.nx, g.ny, g.nz, s->p.ly, s->p.lz, s->p.lxi,
                             s->p.hill_alpha, s->p.flat, g.minsize, s->p.xi_near_wall));
    if (sizeof(C) == 8) {
        s->xiw = d64;
        d64 = nullptr;
    } else {
        HIPCHK(hipMalloc(&s->xiw, nxi * sizeof(float)));
        const long long ng = (long long)(nxi / 7);
        hipLaunchKernelGGL(k_round_weights7, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s->stream, d64,
                           (float*)s->xiw, ng);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    for (double* q : {d64, dzc, dy, dxi})
        if (q) HIPCHK(hipFree(q));

    DevTables<C> t;
    t.xw0 = (const C*)s->xw0; t.xw2 = (const C*)s->xw2; t.yw0 = (const C*)s->yw0; t.yw2 = (const C*)s->yw2;
    t.xiw = (const C*)s->xiw; t.bflag = s->bflag; t.bq = (const C*)s->bq; t.bw = (const C*)s->bw;
    t.les_delta = (const C*)s->les;
    if constexpr (sizeof(C) == 8)
        s->td = t;
    else
        s->tf = t;
    return ISLBM_OK;
}

// host FP64 view [19][n] <-> device storage of one population buffer [19][qs]
hipError_t copy_pops(islbm_solver* s, void* dst, const void* src, size_t vb, hipMemcpyKind kind)
{
    const size_t row = (size_t)s->n * vb, pitch = (size_t)s->qs * vb;
    if (kind == hipMemcpyHostToDevice) return hipMemcpy2D(dst, pitch, src, row, row, NQ, kind);
    return hipMemcpy2D(dst, row, src, pitch, row, NQ, kind);
}

int put_f(islbm_solver* s, int buf, const double* f)
{
    const long long n = s->n;
    const size_t tot = (size_t)NQ * n;
    if (s->store == Store::F64) {
        HIPCHK(copy_pops(s, s->f[buf], f, 8, hipMemcpyHostToDevice));
        return ISLBM_OK;
    }
    if (s->store == Store::F32) {
        std::vector<float> h(tot);
        for (int q = 0; q < NQ; ++q)
            for (long long t = 0; t < n; ++t) h[q * n + t] = (float)(f[q * n + t] - wq(q));
        HIPCHK(copy_pops(s, s->f[buf], h.data(), 4, hipMemcpyHostToDevice));
        return ISLBM_OK;
    }
    std::vector<uint16_t> h(tot);
    for (int q = 0; q < NQ; ++q) {
        const double sc = s->scale[q];
        for (long long t = 0; t < n; ++t) {
            const float v = (float)((f[q * n + t] - wq(q)) * sc);
            if (s->store == Store::BF16) {
                h[q * n + t] = f32_to_bf16_rne(v);
            } else {
                _Float16 hv = (_Float16)v;
                std::memcpy(&h[q * n + t], &hv, 2);
            }
        }
    }
    HIPCHK(copy_pops(s, s->f[buf], h.data(), 2, hipMemcpyHostToDevice));
    return ISLBM_OK;
}

int get_f(islbm_solver* s, int buf, double* f)
{
    const long long n = s->n;
    const size_t tot = (size_t)NQ * n;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->store == Store::F64) {
        HIPCHK(copy_pops(s, f, s->f[buf], 8, hipMemcpyDeviceToHost));
        return ISLBM_OK;
    }
    if (s->store == Store::F32) {
        std::vector<float> h(tot);
        HIPCHK(copy_pops(s, h.data(), s->f[buf], 4, hipMemcpyDeviceToHost));
        for (int q = 0; q < NQ; ++q)
            for (long long t = 0; t < n; ++t) f[q * n + t] = (double)h[q * n + t] + wq(q);
        return ISLBM_OK;
    }
    std::vector<uint16_t> h(tot);
    HIPCHK(copy_pops(s, h.data(), s->f[buf], 2, hipMemcpyDeviceToHost));
    for (int q = 0; q < NQ; ++q) {
        const double isc = 1.0 / (double)s->scale[q];
        for (long long t = 0; t < n; ++t) {
            float v;
            if (s->store == Store::BF16) {
                v = bf16_to_f32(h[q * n + t]);
            } else {
                _Float16 hv;
                std::memcpy(&hv, &h[q * n + t], 2);
                v = (float)hv;
            }
            f[q * n + t] = (double)v * isc + wq(q);
        }
    }
    return ISLBM_OK;
}

StepLaunch make_launch(islbm_solver* s, int cur)
{
    StepLaunch L{};
    L.store = s->store;
    L.fp64 = s->fp64;
    L.kernel = s->kernel;
    L.g = LaunchGeom{s->p.nx, s->p.ny, s->p.nz, s->n, s->qs};
    L.fo = s->f[cur];
    L.fn = s->f[1 - cur];
    L.rho = s->mac[0];
    L.u = s->mac[1];
    L.v = s->mac[2];
    L.w = s->mac[3];
    L.hbuf = s->hbuf;
    L.tables = s->fp64 ? (const void*)&s->td : (const void*)&s->tf;
    L.force = s->scal + 0;
    L.rho_modify = s->scal + 1;
    L.tau = s->g.tau;
    L.dt = s->g.dt;
    L.cs_smag = s->p.cs_smag;
    L.scale = (s->store == Store::BF16 || s->store == Store::FP16) ? s->scale : nullptr;
    L.flux_part = s->flux_part;
    L.flux_dx = s->flux_dxz;
    L.flux_dz = s->flux_dxz ? s->flux_dxz + s->p.nx : nullptr;
    L.flux_acc = s->scal + 2;
    L.flux_plane = s->g.flux_plane;
    L.flux_par = cur;
    L.flux_nslot = s->flux_nslot;
    L.wide_addr = s->wide_addr;
    L.step_ctr = (s->store == Store::BF16 || s->store == Store::FP16) ? s->scal + 3 : nullptr;
    return L;
}

// After v was set from outside the step: the probe partials the next fused step adds are
// those of the new v (the reference's probe reads whatever v holds before the step).
int flux_refresh(islbm_solver* s)
{
    if (!s->flux_part) return ISLBM_OK;
    HIPCHK(launch_flux_refresh(s->stream, s->mac[2], s->fp64, s->flux_dxz, s->flux_dxz + s->p.nx,
                               s->flux_part + (size_t)(1 - s->cur) * s->flux_nslot, s->flux_nslot, s->p.nx, s->p.ny,
                               s->p.nz, s->g.flux_plane));
    return ISLBM_OK;
}

int enqueue_force_update(islbm_solver* s, hipStream_t st = nullptr)
{
    const double beta = (0.001 > 3.0 / s->p.re) ? 0.001 : 3.0 / s->p.re;
    HIPCHK(launch_force_update(st ? st : s->stream, s->scal + 0, s->scal + 2,
                               1.0 / (double)(s->p.lx * (s->p.lz - 1.0)), s->p.ndtfrc, beta, s->g.uref));
    return ISLBM_OK;
}

// one Launch_CollisionStreaming; ev (optional) brackets the stream-collide launch
int enqueue_step(islbm_solver* s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (!s->flux_part)  // the tiled kernel probes the flux itself
        HIPCHK(launch_flux_probe(s->stream, s->mac[2], s->fp64, s->fluxw, s->fluxw + (size_t)s->p.nx * s->p.nz,
                                 s->scal + 2, s->p.nx, s->p.ny, s->p.nz, s->g.flux_plane));
    if (ev0) HIPCHK(hipEventRecord(ev0, s->stream));
    StepLaunch L = make_launch(s, s->cur);
    HIPCHK(launch_stream_collide(s->stream, L));
    if (ev1) HIPCHK(hipEventRecord(ev1, s->stream));
    s->cur = 1 - s->cur;
    s->steps++;
    if (s->st_acc && s->st_every > 0 && s->steps % s->st_every == 0) {
        HIPCHK(launch_stats(s->stream, s->fp64, s->mac[0], s->st_acc, s->st_zdw, s->g.dx, s->g.dy, s->p.nx, s->p.ny,
                            s->p.nz));
        s->st_count++;
    }
    if (s->auto_force && s->steps % s->p.ndtfrc == 0) return enqueue_force_update(s);
    return ISLBM_OK;
}

}  // namespace

extern "C" {

int islbm_default_params(islbm_params* p)
{
    if (!p) return fail(ISLBM_ERR_ARG, "null params");
    std::memset(p, 0, sizeof *p);
    p->nx = 32; p->ny = 128; p->nz = 64;              // variables.h:10-12 (global NY)
    p->lx = 4.5; p->ly = 9.0; p->lz = 3.036;          // variables.h:6-8
    p->cfl = 0.6;                                      // variables.h:22
    p->lxi = 10.0;                                     // variables.h:30
    p->hill_alpha = 1.0;
    p->flat = 0;
    p->mesh_mode = ISLBM_MESH_BISECT;
    p->stretch = 2.0;
    p->tau = 0.6833;                                   // variables.h:95
    p->re = 50.0;                                      // variables.h:41
    p->u_ref = 0.0;
    p->cs_smag = 0.0;
    p->ndtfrc = 10000;                                 // variables.h:48
    p->precision = ISLBM_FP64;
    p->kernel = ISLBM_KERNEL_AUTO;
    p->device = 0;
    return ISLBM_OK;
}

int islbm_required_bytes(const islbm_params* p, size_t* bytes)
{
    if (!p || !bytes) return fail(ISLBM_ERR_ARG, "null argument");
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0) return fail(ISLBM_ERR_ARG, "non-positive grid");
    *bytes = state_bytes(*p);
    return ISLBM_OK;
}

int islbm_create(const islbm_params* pin, size_t budget, islbm_solver** out)
{
    if (!pin || !out) return fail(ISLBM_ERR_ARG, "null argument");
    *out = nullptr;
    if (pin->precision < ISLBM_FP64 || pin->precision > ISLBM_FP16) return fail(ISLBM_ERR_ARG, "unknown precision");
    if (pin->kernel < ISLBM_KERNEL_AUTO || pin->kernel > ISLBM_KERNEL_TILED_WIDE) return fail(ISLBM_ERR_ARG, "unknown kernel");
    islbm_solver* s = new islbm_solver();
    s->p = *pin;
    std::string msg = build_host_geometry(s->p, s->g);
    if (!msg.empty()) {
        delete s;
        return fail(ISLBM_ERR_ARG, "%s", msg.c_str());
    }
    s->bytes = state_bytes(s->p);
    if (budget && s->bytes > budget) {
        delete s;
        return fail(ISLBM_ERR_BUDGET, "solve needs %zu bytes, budget %zu", s->bytes, budget);
    }
    s->n = (long long)s->p.nx * s->p.ny * s->p.nz;
    s->qs = pop_stride(s->n, (int)store_bytes(s->p.precision));
    s->store = store_of(s->p.precision);
    s->fp64 = s->p.precision == ISLBM_FP64;
    for (int q = 0; q < NQ; ++q) s->scale[q] = dev_scale(s->p.precision, q);
    const LaunchGeom lg{s->p.nx, s->p.ny, s->p.nz, s->n, s->qs};
    if (s->p.kernel == ISLBM_KERNEL_AUTO)
        s->kernel = s->fp64 ? ISLBM_KERNEL_TWOPASS
                            : (tiled_supported(lg, s->store, s->fp64) ? ISLBM_KERNEL_TILED : ISLBM_KERNEL_PERNODE);
    else
        s->kernel = s->p.kernel == ISLBM_KERNEL_TILED_WIDE ? ISLBM_KERNEL_TILED : s->p.kernel;
    // wide addressing: asked for, or needed by the grid
    s->wide_addr = s->kernel == ISLBM_KERNEL_TILED &&
                   (s->p.kernel == ISLBM_KERNEL_TILED_WIDE || tiled_needs_wide(lg, s->store));
    if (s->kernel == ISLBM_KERNEL_TILED && !tiled_supported(lg, s->store, s->fp64)) {
        delete s;
        return fail(ISLBM_ERR_UNSUPPORTED, "tiled kernel does not support this grid/precision");
    }
    if (s->kernel == ISLBM_KERNEL_TWOPASS && !s->fp64) {
        delete s;
        return fail(ISLBM_ERR_UNSUPPORTED, "the two-pass kernel computes in FP64 only");
    }
    auto bail = [&](int rc) {
        free_all(s);
        delete s;
        return rc;
    };
    if (hipSetDevice(s->p.device) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "hipSetDevice(%d) failed", s->p.device));
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(ISLBM_ERR_HIP, "hipStreamCreate failed"));
    const size_t fb = (size_t)NQ * s->qs * store_bytes(s->p.precision);
    const size_t mb = (size_t)s->n * compute_bytes(s->p.precision);
    for (int b = 0; b < 2; ++b) {
        if (hipMalloc(&s->f[b], fb) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "hipMalloc f (%zu bytes)", fb));
        if (hipMemset(s->f[b], 0, fb) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "hipMemset f"));
    }
    if (s->kernel == ISLBM_KERNEL_TWOPASS) {
        const size_t hb = (size_t)16 * s->qs * sizeof(double);
        if (hipMalloc(&s->hbuf, hb) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "hipMalloc filtered (%zu bytes)", hb));
    }
    // rho, u, v, w in one allocation, [4][nz][ny][nx]
    if (hipMalloc(&s->mac[0], 4 * mb) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "hipMalloc macros"));
    if (hipMemset(s->mac[0], 0, 4 * mb) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "hipMemset macros"));
    for (int a = 1; a < 4; ++a) s->mac[a] = (char*)s->mac[0] + a * mb;
    if (hipMalloc((void**)&s->scal, 8 * sizeof(double)) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "hipMalloc scalars"));
    if (hipMalloc((void**)&s->red, (kSumBlocks + 1) * sizeof(double)) != hipSuccess)
        return bail(fail(ISLBM_ERR_HIP, "hipMalloc reduction buffer"));
    const double sc[8] = {s->g.force0, 0.0, 0.0, 0, 0, 0, 0, 0};
    if (hipMemcpy(s->scal, sc, sizeof sc, hipMemcpyHostToDevice) != hipSuccess) return bail(fail(ISLBM_ERR_HIP, "scalars"));
    int rc = s->fp64 ? upload_tables<double>(s) : upload_tables<float>(s);
    if (rc) return bail(rc);
    if (s->kernel == ISLBM_KERNEL_TILED) {
        s->flux_nslot = tiled_flux_slots(lg, s->store);
        const size_t pb = 2 * (size_t)s->flux_nslot * sizeof(double);
        std::vector<double> dxz(s->g.flux_dx);
        dxz.insert(dxz.end(), s->g.flux_dz.begin(), s->g.flux_dz.end());
        if (hipMalloc((void**)&s->flux_part, pb) != hipSuccess || hipMemset(s->flux_part, 0, pb) != hipSuccess ||
            hipMalloc((void**)&s->flux_dxz, dxz.size() * sizeof(double)) != hipSuccess ||
            hipMemcpy(s->flux_dxz, dxz.data(), dxz.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(ISLBM_ERR_HIP, "flux probe buffers"));
    }
    *out = s;
    return ISLBM_OK;
}

int islbm_destroy(islbm_solver* s)
{
    if (!s) return ISLBM_OK;
    (void)hipSetDevice(s->p.device);
    (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return ISLBM_OK;
}

int islbm_init_equilibrium(islbm_solver* s, double rho0, const double u0[3], double amp, uint64_t seed)
{
    if (!s || !u0) return fail(ISLBM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->p.device));
    // computed on the device (k_init_equilibrium): no host arrays, no upload
    s->cur = 0;
    HIPCHK(launch_init_equilibrium(s->stream, s->store, s->f[0], s->qs, s->mac[0], s->n, s->p.nx, s->p.ny, s->p.nz,
                                   rho0, u0, amp, seed, s->scale));
    // force, rho_modify, flux accumulator, and the step counters of 16-bit storage
    const double sc[5] = {s->g.force0, 0.0, 0.0, 0.0, 0.0};
    HIPCHK(hipMemcpyAsync(s->scal, sc, sizeof sc, hipMemcpyHostToDevice, s->stream));
    if (int rc = flux_refresh(s)) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->steps = 0;
    return ISLBM_OK;
}

int islbm_upload_f(islbm_solver* s, const double* f)
{
    if (!s || !f) return fail(ISLBM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    return put_f(s, s->cur, f);
}

int islbm_download_f(islbm_solver* s, double* f)
{
    if (!s || !f) return fail(ISLBM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->p.device));
    return get_f(s, s->cur, f);
}

int islbm_download_macros(islbm_solver* s, double* rho, double* u, double* v, double* w)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    double* outs[4] = {rho, u, v, w};
    std::vector<float> h;
    for (int a = 0; a < 4; ++a) {
        if (!outs[a]) continue;
        if (s->fp64) {
            HIPCHK(hipMemcpy(outs[a], s->mac[a], s->n * 8, hipMemcpyDeviceToHost));
        } else {
            h.resize(s->n);
            HIPCHK(hipMemcpy(h.data(), s->mac[a], s->n * 4, hipMemcpyDeviceToHost));
            for (long long t = 0; t < s->n; ++t) outs[a][t] = h[t];
        }
    }
    return ISLBM_OK;
}

int islbm_step(islbm_solver* s, int32_t nsteps)
{
    if (!s || nsteps < 0) return fail(ISLBM_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(s->p.device));
    for (int t = 0; t < nsteps; ++t) {
        int rc = enqueue_step(s, nullptr, nullptr);
        if (rc) return rc;
    }
    return ISLBM_OK;
}

int islbm_step_graph(islbm_solver* s, int32_t nsteps)
{
    if (!s || nsteps <= 0 || nsteps % 2) return fail(ISLBM_ERR_ARG, "islbm_step_graph: nsteps must be even and positive");
    HIPCHK(hipSetDevice(s->p.device));
    // a forcing update inside the chunk is tied to the absolute step count, and so is the
    // statistics cadence: step eagerly
    if (s->auto_force && (s->steps + nsteps) / s->p.ndtfrc > s->steps / s->p.ndtfrc) return islbm_step(s, nsteps);
    if (s->st_acc && s->st_every > 0) return islbm_step(s, nsteps);
    // the chunk recorded from this buffer parity (after an even number of steps the parity
    // returns to cur, so a graph is replayed only from the parity it was recorded at)
    hipGraphExec_t& gx = s->graph[s->cur];
    if (!gx || s->graph_steps[s->cur] != nsteps) {
        if (gx) {
            // a chunk launched from it may still run on the stream
            HIPCHK(hipStreamSynchronize(s->stream));
            HIPCHK(hipGraphExecDestroy(gx));
            gx = nullptr;
            s->graph_steps[s->cur] = 0;
        }
        // record the chunk without running it
        const long long steps0 = s->steps;
        const int cur0 = s->cur;
        const bool af = s->auto_force;
        s->auto_force = false;
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = ISLBM_OK;
        for (int t = 0; t < nsteps && rc == ISLBM_OK; ++t) rc = enqueue_step(s, nullptr, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->stream, &g);
        s->auto_force = af;
        s->steps = steps0;
        s->cur = cur0;
        if (rc != ISLBM_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(ISLBM_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            gx = nullptr;
            return fail(ISLBM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ei));
        }
        s->graph_steps[s->cur] = nsteps;
    }
    HIPCHK(hipGraphLaunch(gx, s->stream));
    s->steps += nsteps;
    return ISLBM_OK;
}

int islbm_stream(islbm_solver* s, void** stream)
{
    if (!s || !stream) return fail(ISLBM_ERR_ARG, "null argument");
    *stream = (void*)s->stream;
    return ISLBM_OK;
}

int islbm_device_bytes(islbm_solver* s, size_t* bytes)
{
    if (!s || !bytes) return fail(ISLBM_ERR_ARG, "null argument");
    *bytes = s->bytes;
    return ISLBM_OK;
}

int islbm_upload_macros(islbm_solver* s, const double* rho, const double* u, const double* v, const double* w)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    const double* in[4] = {rho, u, v, w};
    std::vector<float> h;
    for (int a = 0; a < 4; ++a) {
        if (!in[a]) continue;
        if (s->fp64) {
            HIPCHK(hipMemcpy(s->mac[a], in[a], s->n * 8, hipMemcpyHostToDevice));
        } else {
            h.assign(in[a], in[a] + s->n);
            HIPCHK(hipMemcpy(s->mac[a], h.data(), s->n * 4, hipMemcpyHostToDevice));
        }
    }
    if (v) {
        if (int rc = flux_refresh(s)) return rc;
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return ISLBM_OK;
}

int islbm_mass_correct(islbm_solver* s, double rho0, double* mean_rho)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(launch_sum(s->stream, s->fp64, s->mac[0], s->n, s->red, s->red + kSumBlocks));
    double sum = 0.0;
    HIPCHK(hipMemcpyAsync(&sum, s->red + kSumBlocks, sizeof sum, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const double mean = sum / (double)s->n;
    const double dr = rho0 - mean;
    HIPCHK(hipMemcpy(s->scal + 1, &dr, sizeof dr, hipMemcpyHostToDevice));
    if (mean_rho) *mean_rho = mean;
    return ISLBM_OK;
}

int islbm_mass_correct_noalloc(islbm_solver* s, double rho0, double* mean_rho)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    double* scr = nullptr;
    HIPCHK(sum_scratch(s->p.device, &scr));
    // the reduction fills all kSumBlocks partials and the sum whatever the grid size, so
    // every solver fits the scratch and nothing in it has to start at zero
    static_assert(kSumBlocks + 1 <= kSumScratch, "reduction scratch");
    HIPCHK(launch_sum(s->stream, s->fp64, s->mac[0], s->n, scr, scr + kSumBlocks));
    double sum = 0.0;
    HIPCHK(hipMemcpyAsync(&sum, scr + kSumBlocks, sizeof sum, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const double mean = sum / (double)s->n;
    const double dr = rho0 - mean;
    HIPCHK(hipMemcpy(s->scal + 1, &dr, sizeof dr, hipMemcpyHostToDevice));
    if (mean_rho) *mean_rho = mean;
    return ISLBM_OK;
}

int islbm_monitor(islbm_solver* s, double out[4])
{
    if (!s || !out) return fail(ISLBM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    // the reference's monitor node (NX6/2, NYD6/2, NZ6/2) of the padded single-rank layout
    const int i = (s->p.nx + 7) / 2 - 3, j = (s->p.ny + 7) / 2 - 3, k = (s->p.nz + 6) / 2 - 3;
    const long long id = ((long long)k * s->p.ny + j) * s->p.nx + i;
    const int order[3] = {2, 1, 3};  // v, u, w (monitor.h:19)
    for (int t = 0; t < 3; ++t) {
        if (s->fp64) {
            HIPCHK(hipMemcpy(&out[t], (const double*)s->mac[order[t]] + id, 8, hipMemcpyDeviceToHost));
        } else {
            float f = 0.f;
            HIPCHK(hipMemcpy(&f, (const float*)s->mac[order[t]] + id, 4, hipMemcpyDeviceToHost));
            out[t] = f;
        }
    }
    HIPCHK(hipMemcpy(&out[3], s->scal + 0, 8, hipMemcpyDeviceToHost));
    return ISLBM_OK;
}

int islbm_stats_enable(islbm_solver* s, int32_t every)
{
    if (!s || every < 0) return fail(ISLBM_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    const size_t nacc = (size_t)NSTAT * s->n;
    if (!s->st_acc) {
        HIPCHK(hipMalloc((void**)&s->st_acc, nacc * sizeof(double)));
        const std::vector<double> zw = zderiv_weights(s->g);
        HIPCHK(hipMalloc((void**)&s->st_zdw, zw.size() * sizeof(double)));
        HIPCHK(hipMemcpy(s->st_zdw, zw.data(), zw.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemset(s->st_acc, 0, nacc * sizeof(double)));
    s->st_every = every;
    s->st_count = 0;
    return ISLBM_OK;
}

int islbm_stats_disable(islbm_solver* s)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (double** q : {&s->st_acc, &s->st_zdw})
        if (*q) {
            HIPCHK(hipFree(*q));
            *q = nullptr;
        }
    s->st_every = 0;
    s->st_count = 0;
    return ISLBM_OK;
}

int islbm_stats_accumulate(islbm_solver* s)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    if (!s->st_acc) return fail(ISLBM_ERR_STATE, "statistics are not enabled");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(launch_stats(s->stream, s->fp64, s->mac[0], s->st_acc, s->st_zdw, s->g.dx, s->g.dy, s->p.nx, s->p.ny,
                        s->p.nz));
    s->st_count++;
    return ISLBM_OK;
}

int islbm_stats_reset(islbm_solver* s)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    if (!s->st_acc) return fail(ISLBM_ERR_STATE, "statistics are not enabled");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemset(s->st_acc, 0, (size_t)NSTAT * s->n * sizeof(double)));
    s->st_count = 0;
    return ISLBM_OK;
}

int islbm_stats_count(islbm_solver* s, int64_t* count)
{
    if (!s || !count) return fail(ISLBM_ERR_ARG, "null argument");
    *count = s->st_count;
    return ISLBM_OK;
}

int islbm_stats_set_count(islbm_solver* s, int64_t count)
{
    if (!s || count < 0) return fail(ISLBM_ERR_ARG, "bad argument");
    s->st_count = count;
    return ISLBM_OK;
}

namespace {
int stat_index(const char* name)
{
    for (int a = 0; a < NSTAT; ++a)
        if (std::strcmp(name, kStatNames[a]) == 0) return a;
    return -1;
}
}  // namespace

int islbm_stats_names(const char** names, int32_t cap)
{
    if (!names || cap < NSTAT) return -NSTAT;
    for (int a = 0; a < NSTAT; ++a) names[a] = kStatNames[a];
    return NSTAT;
}

int islbm_stats_download(islbm_solver* s, const char* name, double* out)
{
    if (!s || !name || !out) return fail(ISLBM_ERR_ARG, "null argument");
    if (!s->st_acc) return fail(ISLBM_ERR_STATE, "statistics are not enabled");
    const int a = stat_index(name);
    if (a < 0) return fail(ISLBM_ERR_ARG, "unknown statistic %s", name);
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(out, s->st_acc + (size_t)a * s->n, s->n * sizeof(double), hipMemcpyDeviceToHost));
    return ISLBM_OK;
}

int islbm_stats_upload(islbm_solver* s, const char* name, const double* in)
{
    if (!s || !name || !in) return fail(ISLBM_ERR_ARG, "null argument");
    if (!s->st_acc) return fail(ISLBM_ERR_STATE, "statistics are not enabled");
    const int a = stat_index(name);
    if (a < 0) return fail(ISLBM_ERR_ARG, "unknown statistic %s", name);
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(s->st_acc + (size_t)a * s->n, in, s->n * sizeof(double), hipMemcpyHostToDevice));
    return ISLBM_OK;
}

int islbm_sync(islbm_solver* s)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ISLBM_OK;
}

int islbm_modify_forcing(islbm_solver* s)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    return enqueue_force_update(s);
}

int islbm_get_force(islbm_solver* s, double* force)
{
    if (!s || !force) return fail(ISLBM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(force, s->scal + 0, sizeof(double), hipMemcpyDeviceToHost));
    return ISLBM_OK;
}

int islbm_set_force(islbm_solver* s, double force)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(s->scal + 0, &force, sizeof(double), hipMemcpyHostToDevice));
    return ISLBM_OK;
}

int islbm_set_rho_modify(islbm_solver* s, double drho)
{
    if (!s) return fail(ISLBM_ERR_ARG, "null solver");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(s->scal + 1, &drho, sizeof(double), hipMemcpyHostToDevice));
    return ISLBM_OK;
}

int islbm_get_flux_sum(islbm_solver* s, double* sum)
{
    if (!s || !sum) return fail(ISLBM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(sum, s->scal + 2, sizeof(double), hipMemcpyDeviceToHost));
    return ISLBM_OK;
}

int islbm_scalars(islbm_solver* s, double out[8])
{
    if (!s || !out) return fail(ISLBM_ERR_ARG, "null argument");
    const HostGeom& g = s->g;
    const double v[8] = {g.dt, g.tau, g.nu, g.uref, g.a, g.dx, g.dy, g.force0};
    std::memcpy(out, v, sizeof v);
    return ISLBM_OK;
}

int islbm_get_params(islbm_solver* s, islbm_params* p)
{
    if (!s || !p) return fail(ISLBM_ERR_ARG, "null argument");
    *p = s->p;
    return ISLBM_OK;
}

int islbm_step_count(islbm_solver* s, int64_t* steps)
{
    if (!s || !steps) return fail(ISLBM_ERR_ARG, "null argument");
    *steps = s->steps;
    return ISLBM_OK;
}

int islbm_set_step_count(islbm_solver* s, int64_t steps)
{
    if (!s || steps < 0) return fail(ISLBM_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(hipStreamSynchronize(s->stream));
    // the device copy read by the 16-bit step, in both parity slots
    const double c[2] = {(double)steps, (double)steps};
    HIPCHK(hipMemcpy(s->scal + 3, c, sizeof c, hipMemcpyHostToDevice));
    s->steps = steps;
    return ISLBM_OK;
}

int islbm_kernel_in_use(islbm_solver* s, int32_t* kernel)
{
    if (!s || !kernel) return fail(ISLBM_ERR_ARG, "null argument");
    *kernel = s->kernel == ISLBM_KERNEL_TILED && s->wide_addr ? ISLBM_KERNEL_TILED_WIDE : s->kernel;
    return ISLBM_OK;
}

int islbm_tile_shape(islbm_solver* s, int32_t* cols, int32_t* rows)
{
    if (!s || !cols || !rows) return fail(ISLBM_ERR_ARG, "null argument");
    *cols = *rows = 0;
    if (s->kernel == ISLBM_KERNEL_TILED) {
        const LaunchGeom lg{s->p.nx, s->p.ny, s->p.nz, s->n, s->qs};
        int c = 0, r = 0;
        tiled_tile_shape(lg, s->store, &c, &r);
        *cols = c;
        *rows = r;
    }
    return ISLBM_OK;
}

int islbm_get_table(islbm_solver* s, const char* name, double* out, int32_t cap)
{
    if (!s || !name || !out) return fail(ISLBM_ERR_ARG, "null argument");
    const HostGeom& g = s->g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    std::vector<double> v;
    const std::string nm(name);
    if (nm == "x") v = g.x;
    else if (nm == "y") v = g.y;
    else if (nm == "xi") v = g.xi;
    else if (nm == "z") v = g.zc;
    else if (nm == "xw0") v = g.xw0;
    else if (nm == "xw2") v = g.xw2;
    else if (nm == "yw0") v = g.yw0;
    else if (nm == "yw2") v = g.yw2;
    else if (nm.rfind("xiw", 0) == 0) {
        static const char* fam[NXI] = {"xiw3", "xiw4", "xiw5", "xiw6", "xiw15", "xiw16", "xiw17", "xiw18"};
        int f = -1;
        for (int t = 0; t < NXI; ++t)
            if (nm == fam[t]) f = t;
        if (f < 0) return fail(ISLBM_ERR_ARG, "unknown table %s", name);
        const size_t nxi = (size_t)nz * ny * NXI * 7;
        std::vector<double> all(nxi);
        HIPCHK(hipSetDevice(s->p.device));
        HIPCHK(hipStreamSynchronize(s->stream));
        if (s->fp64) {
            HIPCHK(hipMemcpy(all.data(), s->xiw, nxi * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> h(nxi);
            HIPCHK(hipMemcpy(h.data(), s->xiw, nxi * 4, hipMemcpyDeviceToHost));
            for (size_t t = 0; t < nxi; ++t) all[t] = h[t];
        }
        v.resize((size_t)nz * ny * 7);
        for (size_t kj = 0; kj < (size_t)nz * ny; ++kj)
            for (int m = 0; m < 7; ++m) v[kj * 7 + m] = all[(kj * NXI + f) * 7 + m];
    } else if (nm.rfind("bfl_flag", 0) == 0 || nm.rfind("bfl_q", 0) == 0) {
        const bool isq = nm.rfind("bfl_q", 0) == 0;
        const std::string suf = nm.substr(isq ? 5 : 8);
        const int b = suf == "3" ? B3 : suf == "4" ? B4 : suf == "15" ? B15 : suf == "16" ? B16 : -1;
        if (b < 0) return fail(ISLBM_ERR_ARG, "unknown table %s", name);
        const size_t nb = 2 * (size_t)ny;
        v.resize(nb);
        for (size_t t = 0; t < nb; ++t) v[t] = isq ? g.bq[b * nb + t] : (double)g.bflag[b * nb + t];
    } else if (nm.rfind("bfl_", 0) == 0) {
        static const char* wn[NBW] = {"bfl_x37", "bfl_x38", "bfl_y378", "bfl_xi378", "bfl_x49", "bfl_x410",
                                      "bfl_y4910", "bfl_xi4910", "bfl_y15", "bfl_xi15", "bfl_y16", "bfl_xi16"};
        int b = -1;
        for (int t = 0; t < NBW; ++t)
            if (nm == wn[t]) b = t;
        if (b < 0) return fail(ISLBM_ERR_ARG, "unknown table %s", name);
        const size_t cnt = 2 * (size_t)ny * 7;
        v.assign(g.bw.begin() + b * cnt, g.bw.begin() + (b + 1) * cnt);
    } else if (nm == "fluxw") {
        v = g.fluxw;
    } else {
        return fail(ISLBM_ERR_ARG, "unknown table %s", name);
    }
    (void)nx;
    if ((int64_t)v.size() > cap) return -(int)v.size();
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return (int)v.size();
}

int islbm_time_steps(islbm_solver* s, int32_t nsteps, double out[2])
{
    if (!s || !out || nsteps <= 0) return fail(ISLBM_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(s->p.device));
    const size_t need = 2 * (size_t)nsteps + 2;
    while (s->evpool.size() < need) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        s->evpool.push_back(e);
    }
    hipEvent_t* ev = s->evpool.data();
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipEventRecord(ev[0], s->stream));
    for (int t = 0; t < nsteps; ++t) {
        int rc = enqueue_step(s, ev[2 + 2 * t], ev[3 + 2 * t]);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(ev[1], s->stream));
    HIPCHK(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    out[0] = ms;
    double ksum = 0.0;
    for (int t = 0; t < nsteps; ++t) {
        float k = 0.f;
        HIPCHK(hipEventElapsedTime(&k, ev[2 + 2 * t], ev[3 + 2 * t]));
        ksum += k;
    }
    out[1] = ksum;
    return ISLBM_OK;
}

const char* islbm_last_error(void) { return g_err.c_str(); }

int islbm_copy_bandwidth(int32_t device, size_t bytes, int32_t iters, double* gbps)
{
    if (!gbps || iters <= 0 || bytes < 16) return fail(ISLBM_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(device));
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    };
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, bytes, st) != hipSuccess) {
        cleanup();
        return fail(ISLBM_ERR_HIP, "copy probe set-up failed");
    }
    hipError_t e = launch_stream_copy(st, a, b, (long long)bytes);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int t = 0; t < iters && e == hipSuccess; ++t) e = launch_stream_copy(st, a, b, (long long)bytes);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    cleanup();
    if (e != hipSuccess) return fail(ISLBM_ERR_HIP, "copy probe: %s", hipGetErrorString(e));
    *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
    return ISLBM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// Grouped stepping of several solves (the serving loop's batches)
// ------------------------------------------------------------------------------------
bool islbm::group_eligible(islbm_solver* s)
{
    const LaunchGeom lg{s->p.nx, s->p.ny, s->p.nz, s->n, s->qs};
    return s->kernel == ISLBM_KERNEL_TILED && !s->wide_addr && tiled_groupable(lg, s->store) &&
           !(s->st_acc && s->st_every > 0) && s->auto_force;
}

int islbm::group_limit(islbm_solver* s, int64_t want)
{
    const int64_t to_force = s->p.ndtfrc - s->steps % s->p.ndtfrc;  // steps up to the next controller update
    return (int)std::min<int64_t>(want, to_force);
}

int islbm::group_parity(islbm_solver* s) { return s->cur; }

int islbm::field_summary(islbm_solver* s, double* part, double out[3])
{
    HIPCHK(hipSetDevice(s->p.device));
    HIPCHK(launch_field_summary(s->stream, s->fp64, s->mac[0], s->n, part, part + 3 * kSumBlocks));
    HIPCHK(hipMemcpyAsync(out, part + 3 * kSumBlocks, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ISLBM_OK;
}

// Enqueues max(limit) grouped launches on st (member e takes limit[e] steps) and, for a
// member whose chunk ends on a controller step, its forcing update; host state untouched.
int islbm::group_record(hipStream_t st, islbm_solver* const* ss, const int* limit, int n, const GroupFork* fork)
{
    if (n <= 0 || n > kGroupMaxSolves) return fail(ISLBM_ERR_ARG, "group of %d solves", n);
    int steps = 0;
    for (int e = 0; e < n; ++e) steps = std::max(steps, limit[e]);
    StepLaunch L[kGroupMaxSolves];
    for (int t = 0; t < steps; ++t) {
        for (int e = 0; e < n; ++e) L[e] = make_launch(ss[e], ss[e]->cur ^ (t & 1));
        HIPCHK(launch_tiled_group(st, L, limit, n, t, fork));
    }
    for (int e = 0; e < n; ++e)
        if ((ss[e]->steps + limit[e]) % ss[e]->p.ndtfrc == 0)
            if (int rc = enqueue_force_update(ss[e], st)) return rc;
    return ISLBM_OK;
}

// The host state after a recorded chunk: buffer parity and step count of each member.
void islbm::group_commit(islbm_solver* const* ss, const int* limit, int n)
{
    for (int e = 0; e < n; ++e) {
        ss[e]->cur ^= (limit[e] & 1);
        ss[e]->steps += limit[e];
    }
}

// ------------------------------------------------------------------------------------
// Reference-named entry points on the reference's padded FP64 layout
// ------------------------------------------------------------------------------------
namespace {
islbm_solver* g_ref = nullptr;
double* g_ref_mac[4] = {nullptr, nullptr, nullptr, nullptr};  // padded u, v, w, rho_d

void ref_release()
{
    for (double*& p : g_ref_mac)
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
    if (g_ref) {
        islbm_destroy(g_ref);
        g_ref = nullptr;
    }
}
}  // namespace

extern "C" {

int islbm_ref_init(const islbm_params* p)
{
    if (!p) return fail(ISLBM_ERR_ARG, "null params");
    if (p->precision != ISLBM_FP64)
        return fail(ISLBM_ERR_UNSUPPORTED, "the reference-named entry points take FP64 populations");
    ref_release();
    islbm_params q = *p;
    q.kernel = ISLBM_KERNEL_AUTO;  // FP64: the two-pass kernel
    int rc = islbm_create(&q, 0, &g_ref);
    if (rc) return rc;
    g_ref->auto_force = false;  // the reference's main calls Launch_ModifyForcingTerm itself
    const size_t padded = (size_t)(p->nx + 7) * (p->ny + 7) * (p->nz + 6);
    for (double*& m : g_ref_mac) {
        HIPCHK(hipMalloc((void**)&m, padded * sizeof(double)));
        HIPCHK(hipMemset(m, 0, padded * sizeof(double)));
    }
    return ISLBM_OK;
}

int islbm_ref_shutdown(void)
{
    ref_release();
    return ISLBM_OK;
}

int islbm_ref_device_arrays(double** u, double** v, double** w, double** rho, double** force, double** rho_modify)
{
    if (!g_ref) return fail(ISLBM_ERR_STATE, "islbm_ref_init has not been called");
    if (u) *u = g_ref_mac[0];
    if (v) *v = g_ref_mac[1];
    if (w) *w = g_ref_mac[2];
    if (rho) *rho = g_ref_mac[3];
    if (force) *force = g_ref->scal + 0;
    if (rho_modify) *rho_modify = g_ref->scal + 1;
    return ISLBM_OK;
}

// Launch_CollisionStreaming (evolution.h:842-1001): f_old -> f_new for every unique node,
// then the ghost and duplicate layers of f_new and of u, v, w, rho_d are refreshed.
// Errors are recorded for islbm_last_error (the reference exits the process instead).
void Launch_CollisionStreaming(double* f_old[19], double* f_new[19])
{
    islbm_solver* s = g_ref;
    if (!s || !f_old || !f_new) {
        fail(ISLBM_ERR_STATE, "Launch_CollisionStreaming: islbm_ref_init has not been called");
        return;
    }
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const long long n = s->n;
    if (hipSetDevice(s->p.device) != hipSuccess) {
        fail(ISLBM_ERR_HIP, "hipSetDevice failed");
        return;
    }
    PtrSet pad{}, uni{};
    for (int q = 0; q < NQ; ++q) {
        pad.p[q] = f_old[q];
        uni.p[q] = (double*)s->f[s->cur] + (long long)q * s->qs;
    }
    // macroscopic state (u, v, w, rho_d): the flux probe reads v of the previous step
    const int mac_of[4] = {1, 2, 3, 0};  // unique order rho,u,v,w <- padded order u,v,w,rho
    for (int a = 0; a < 4; ++a) {
        pad.p[NQ + a] = g_ref_mac[a];
        uni.p[NQ + a] = (double*)s->mac[mac_of[a]];
    }
    // the step recomputes rho, u, v, w from the populations; of the previous step's fields only
    // v is read (by the flux probe), so only v is gathered with the 19 population arrays
    PtrSet pg = pad, ug = uni;
    pg.p[NQ] = pad.p[NQ + 1];
    ug.p[NQ] = uni.p[NQ + 1];
    hipError_t e = launch_padded_to_unique(s->stream, pg, ug, nx, ny, nz, NQ + 1);
    if (e != hipSuccess) {
        fail(ISLBM_ERR_HIP, "Launch_CollisionStreaming gather: %s", hipGetErrorString(e));
        return;
    }
    if (enqueue_step(s, nullptr, nullptr) != ISLBM_OK) return;
    PtrSet pn{}, un{};
    for (int q = 0; q < NQ; ++q) {
        pn.p[q] = f_new[q];
        un.p[q] = (double*)s->f[s->cur] + (long long)q * s->qs;
    }
    for (int a = 0; a < 4; ++a) {
        pn.p[NQ + a] = g_ref_mac[a];
        un.p[NQ + a] = (double*)s->mac[mac_of[a]];
    }
    e = launch_unique_to_padded(s->stream, un, pn, nx, ny, nz, NQ + 4);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) fail(ISLBM_ERR_HIP, "Launch_CollisionStreaming scatter: %s", hipGetErrorString(e));
}

// Launch_ModifyForcingTerm (evolution.h:1003-1042) on the device; single rank.
void Launch_ModifyForcingTerm(void)
{
    if (!g_ref) {
        fail(ISLBM_ERR_STATE, "Launch_ModifyForcingTerm: islbm_ref_init has not been called");
        return;
    }
    if (hipSetDevice(g_ref->p.device) != hipSuccess || enqueue_force_update(g_ref) != ISLBM_OK) return;
    if (hipStreamSynchronize(g_ref->stream) != hipSuccess) fail(ISLBM_ERR_HIP, "Launch_ModifyForcingTerm sync");
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// Host self-check of the collision: the opposite-pair form used by the kernels against
// the row-by-row MRT form, on seeded near-equilibrium states (FP64 and FP32, with and
// without the Smagorinsky relaxation).  No GPU involved.
// ------------------------------------------------------------------------------------
#include "islbm_step.h"

namespace {
template <bool kDev, class C>
double collision_forms_gap(int les, uint64_t& state)
{
    auto uni = [&]() {  // SplitMix64 -> [-1, 1)
        uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    };
    double worst = 0.0;
    for (int it = 0; it < 500; ++it) {
        C A[NQ], B[NQ];
        const double ux = 0.05 * uni(), uy = 0.05 * uni(), uz = 0.05 * uni(), rho = 1.0 + 0.01 * uni();
        for (int q = 0; q < NQ; ++q) {
            const double eu = cx(q) * ux + cy(q) * uy + cz(q) * uz;
            const double f = wq(q) * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * (ux * ux + uy * uy + uz * uz)) * (1 + 0.01 * uni());
            A[q] = B[q] = (C)(kDev ? f - wq(q) : f);
        }
        C r1, u1, v1, w1, t1, r2, u2, v2, w2, t2;
        collide_node<kDev>(A, (C)1e-5, (C)0.0190875, (C)2e-4, (C)0.05, les, (C)0.6, (C)0.1, r1, u1, v1, w1, t1);
        collide_pairs<kDev>(B, (C)1e-5, (C)0.0190875, (C)2e-4, (C)0.05, les, (C)0.6, (C)0.1, r2, u2, v2, w2, t2,
                            (C)(2 * 0.6 - 1));
        for (int q = 0; q < NQ; ++q) worst = std::max(worst, std::fabs((double)A[q] - (double)B[q]));
    }
    return worst;
}
}  // namespace

extern "C" int islbm_selftest_collision(double out[4])
{
    if (!out) return fail(ISLBM_ERR_ARG, "null argument");
    uint64_t st = 20261203;
    out[0] = collision_forms_gap<false, double>(0, st);
    out[1] = collision_forms_gap<false, double>(1, st);
    out[2] = collision_forms_gap<true, float>(0, st);
    out[3] = collision_forms_gap<true, float>(1, st);
    return ISLBM_OK;
}
