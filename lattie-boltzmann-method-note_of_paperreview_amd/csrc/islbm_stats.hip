// islbm_stats.hip -- turbulence statistics accumulation on the device (Launch_TurbulentSum,
// statistics.h:122-249: MeanVars + MeanDerivatives), the reference's global mass sum and
// the monitor-node read-out helpers.
//
// Accumulators are FP64 sums over samples, [a][k][j][i] over the unique nodes, in the
// order of kStatNames.  x and y derivatives are the reference's 4th-order central
// differences (statistics.h:188-194) with periodic wrapping; the wall-normal derivative
// applies 5-point weights along the node's vertical mesh line on the window the reference
// uses (n = k - 2, clamped to the column, statistics.h:196-199).
#include "islbm_internal.h"

#include <mutex>

namespace islbm {

namespace {

template <typename C>
__global__ __launch_bounds__(256) void k_stats(const C* __restrict__ mac, double* __restrict__ acc,
                                               const double* __restrict__ zdw, double inv12dx, double inv12dy, int nx,
                                               int ny, int nz)
{
    const long long n = (long long)nx * ny * nz;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((long long)nx * ny));
    const C* rho = mac;
    const C* U = mac + n;
    const C* V = mac + 2 * n;
    const C* Wv = mac + 3 * n;
    const double u1 = U[id], v1 = V[id], w1 = Wv[id];
    const double p1 = 1.0 / 3.0 * (double)rho[id] - 1.0 / 3.0;
    double s[NSTAT];
    s[0] = u1;
    s[1] = v1;
    s[2] = w1;
    s[3] = p1;
    s[4] = u1 * u1;
    s[5] = u1 * v1;
    s[6] = u1 * w1;
    s[7] = v1 * v1;
    s[8] = v1 * w1;
    s[9] = w1 * w1;
    s[10] = p1 * u1;
    s[11] = p1 * v1;
    s[12] = p1 * w1;
    s[13] = p1 * p1;
    s[14] = 0.5 * (u1 * u1 + v1 * v1 + w1 * w1);
    // derivatives: d/dx, d/dy 4th-order central (periodic); d/dz 5-point on the column
    const long long row = ((long long)k * ny + j) * nx;
    const int im1 = wrap_index(i - 1, nx), ip1 = wrap_index(i + 1, nx), im2 = wrap_index(i - 2, nx),
              ip2 = wrap_index(i + 2, nx);
    const int jm1 = wrap_index(j - 1, ny), jp1 = wrap_index(j + 1, ny), jm2 = wrap_index(j - 2, ny),
              jp2 = wrap_index(j + 2, ny);
    const long long pl = (long long)k * ny * nx;
    int nb = k - 2;
    nb = nb < 0 ? 0 : (nb > nz - 5 ? nz - 5 : nb);
    const double* cw = zdw + ((long long)j * nz + k) * 5;
    const C* fld[3] = {U, V, Wv};
    double d[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const C* f = fld[c];
        d[c][0] = (8.0 * ((double)f[row + ip1] - (double)f[row + im1]) - ((double)f[row + ip2] - (double)f[row + im2])) *
                  inv12dx;
        d[c][1] = (8.0 * ((double)f[pl + (long long)jp1 * nx + i] - (double)f[pl + (long long)jm1 * nx + i]) -
                   ((double)f[pl + (long long)jp2 * nx + i] - (double)f[pl + (long long)jm2 * nx + i])) *
                  inv12dy;
        double dz = 0.0;
#pragma unroll
        for (int m = 0; m < 5; ++m) dz += (double)f[((long long)(nb + m) * ny + j) * nx + i] * cw[m];
        d[c][2] = dz;
    }
    // DUDX2 DUDY2 DUDZ2 DVDX2 ... DWDZ2
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a) s[15 + 3 * c + a] = d[c][a] * d[c][a];
    s[24] = u1 * u1 * u1;
    s[25] = u1 * u1 * v1;
    s[26] = u1 * u1 * w1;
    s[27] = v1 * v1 * u1;
    s[28] = v1 * v1 * v1;
    s[29] = v1 * v1 * w1;
    s[30] = w1 * w1 * u1;
    s[31] = w1 * w1 * v1;
    s[32] = w1 * w1 * w1;
#pragma unroll
    for (int a = 0; a < NSTAT; ++a) acc[(long long)a * n + id] += s[a];
}

// deterministic two-pass sum of n values: fixed blocks, fixed order
template <typename C>
__global__ __launch_bounds__(256) void k_sum_partial(const C* __restrict__ x, long long n, double* __restrict__ part)
{
    __shared__ double red[256];
    const long long per = (n + gridDim.x - 1) / gridDim.x;
    const long long b0 = (long long)blockIdx.x * per;
    const long long b1 = b0 + per < n ? b0 + per : n;
    double acc = 0.0;
    for (long long t = b0 + threadIdx.x; t < b1; t += blockDim.x) acc += (double)x[t];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_final(const double* __restrict__ part, int np, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = 0.0;
    for (int t = 0; t < np; ++t) acc += part[t];
    out[0] = acc;
}

// sum rho, sum v and max |u| over fixed blocks in fixed order; a NaN speed stays NaN
template <typename C>
__global__ __launch_bounds__(256) void k_summary_partial(const C* __restrict__ mac, long long n, double* __restrict__ part)
{
    __shared__ double red[3][256];
    const long long per = (n + gridDim.x - 1) / gridDim.x;
    const long long b0 = (long long)blockIdx.x * per;
    const long long b1 = b0 + per < n ? b0 + per : n;
    double sr = 0.0, sv = 0.0, um = 0.0;
    for (long long t = b0 + threadIdx.x; t < b1; t += blockDim.x) {
        const double u = (double)mac[n + t], v = (double)mac[2 * n + t], w = (double)mac[3 * n + t];
        sr += (double)mac[t];
        sv += v;
        const double m = sqrt(u * u + v * v + w * w);
        um = (m != m || um != um) ? m + um : (m > um ? m : um);
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = sv;
    red[2][threadIdx.x] = um;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            const double a = red[2][threadIdx.x], c = red[2][threadIdx.x + w];
            red[2][threadIdx.x] = (a != a || c != c) ? a + c : (c > a ? c : a);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[(long long)q * gridDim.x + blockIdx.x] = red[q][0];
}

__global__ void k_summary_final(const double* __restrict__ part, int np, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sr = 0.0, sv = 0.0, um = 0.0;
    for (int t = 0; t < np; ++t) {
        sr += part[t];
        sv += part[np + t];
        const double m = part[2 * np + t];
        um = (m != m || um != um) ? m + um : (m > um ? m : um);
    }
    out[0] = sr;
    out[1] = sv;
    out[2] = um;
}

}  // namespace

const char* const kStatNames[NSTAT] = {"U",     "V",     "W",     "P",     "UU",    "UV",    "UW",   "VV",   "VW",
                                        "WW",    "PU",    "PV",    "PW",    "PP",    "KT",    "DUDX2", "DUDY2", "DUDZ2",
                                        "DVDX2", "DVDY2", "DVDZ2", "DWDX2", "DWDY2", "DWDZ2", "UUU",  "UUV",  "UUW",
                                        "VVU",   "VVV",   "VVW",   "WWU",   "WWV",   "WWW"};

hipError_t launch_stats(hipStream_t st, bool fp64, const void* mac, double* acc, const double* zdw, double dx, double dy,
                        int nx, int ny, int nz)
{
    const long long n = (long long)nx * ny * nz;
    const unsigned nb = (unsigned)((n + 255) / 256);
    const double ix = 1.0 / (12.0 * dx), iy = 1.0 / (12.0 * dy);
    if (fp64)
        hipLaunchKernelGGL(k_stats<double>, dim3(nb), dim3(256), 0, st, (const double*)mac, acc, zdw, ix, iy, nx, ny, nz);
    else
        hipLaunchKernelGGL(k_stats<float>, dim3(nb), dim3(256), 0, st, (const float*)mac, acc, zdw, ix, iy, nx, ny, nz);
    return hipGetLastError();
}

// Scratch of the allocation-free reduction (islbm_mass_correct_noalloc): the kSumBlocks
// partials and the sum, one copy per GPU for the life of the process.  Host code never
// uses this name as an address: the device address comes from sum_scratch() below.
static __device__ __align__(256) double scratch[kSumScratch];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t sum_scratch(int device, double** out)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // the calling thread has `device` current (the entry point sets it)
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = (double*)g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_field_summary(hipStream_t st, bool fp64, const void* mac, long long n, double* part, double* out)
{
    if (fp64)
        hipLaunchKernelGGL(k_summary_partial<double>, dim3(kSumBlocks), dim3(256), 0, st, (const double*)mac, n, part);
    else
        hipLaunchKernelGGL(k_summary_partial<float>, dim3(kSumBlocks), dim3(256), 0, st, (const float*)mac, n, part);
    hipLaunchKernelGGL(k_summary_final, dim3(1), dim3(64), 0, st, (const double*)part, kSumBlocks, out);
    return hipGetLastError();
}

hipError_t launch_sum(hipStream_t st, bool fp64, const void* x, long long n, double* part, double* out)
{
    if (fp64)
        hipLaunchKernelGGL(k_sum_partial<double>, dim3(kSumBlocks), dim3(256), 0, st, (const double*)x, n, part);
    else
        hipLaunchKernelGGL(k_sum_partial<float>, dim3(kSumBlocks), dim3(256), 0, st, (const float*)x, n, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, st, (const double*)part, kSumBlocks, out);
    return hipGetLastError();
}

}  // namespace islbm
