// islbm_kernels.hip -- gfx950 kernels of the ISLBM periodic-hill step.
//
//   k_xi_weights        FP64 per-node xi interpolation weights (GetIntrplParameter_Xi,
//                       initialization.h:185-258), 8 departure families per node.
//   k_collide_pernode   one thread per node: interpolated streaming of all 19 populations
//                       (interpolationHillISLBM.h:4-161, evolution.h:451-475), flat-wall
//                       bounce-back (:482-496), BFL (:505-572), moments, equilibrium and MRT
//                       with body force (:597-679, MRT_Process.h:4-179), optional
//                       Smagorinsky relaxation.  x and y periodicity by index wrap, which
//                       replaces periodicSW (:732-777) and the y halo exchange (:980-994).
//   k_flux_partial/final FP64 bulk-flux sum over the probe plane (AccumulateUbulk, :826-840).
//   k_force_update      forcing controller (Launch_ModifyForcingTerm, :1003-1042).
//
// Population storage is [q][k][j][i].  FP64 stores f itself; FP32 stores f - w_q; the
// 16-bit formats store (f - w_q) * scale_q.  Reduced-precision kernels compute in the
// deviation form throughout (rho = 1 + sum), which keeps FP32 cancellation small.
#include "islbm_internal.h"
#include "islbm_step.h"

namespace islbm {

// ------------------------------------------------------------------------------------
// Set-up kernels
// ------------------------------------------------------------------------------------

__global__ void k_xi_weights(double* __restrict__ out, const double* __restrict__ zc,
                             const double* __restrict__ y, const double* __restrict__ xi, int ny, int nz,
                             double ly, double lz, double lxi, double alpha, int flat, double ms, int near_wall)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nz * ny * NXI) return;
    const int f = t % NXI;
    const int jk = t / NXI;
    const int j = jk % ny;
    const int k = jk / ny;
    const double pz = zc[(size_t)j * nz + k] + xi_dz(f) * ms;
    const double py = y[j] + xi_dy(f) * ms;
    const double H = hill_profile(py, ly, alpha, flat);
    const double L = lz - H - ms;
    const double pxi = lxi * (pz - (H + ms / 2.0)) / L;
    double* o = out + ((size_t)(k * ny + j) * NXI + f) * 7;
    xi_weights7(pxi, xi, k, nz, near_wall, o);
}

hipError_t launch_xi_weights(hipStream_t st, double* out, const double* zc, const double* y, const double* xi,
                             int nx, int ny, int nz, double ly, double lz, double lxi, double alpha, int flat,
                             double minsize, int near_wall)
{
    (void)nx;
    const int n = nz * ny * NXI;
    hipLaunchKernelGGL(k_xi_weights, dim3((n + 255) / 256), dim3(256), 0, st, out, zc, y, xi, ny, nz, ly, lz, lxi,
                       alpha, flat, minsize, near_wall);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Per-node stream-collide
// ------------------------------------------------------------------------------------

template <Store ST>
__global__ __launch_bounds__(256) void k_collide_pernode(StepArgs<ST> a)
{
    using CD = Codec<ST>;
    using C = typename CD::C;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= a.n) return;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i = (int)(gid % nx);
    const long long t = gid / nx;
    const int j = (int)(t % ny);
    const int k = (int)(t / ny);
    const long long plane = (long long)nx * ny;
    const long long N = a.n;

    int xo[7], yo[7];
#pragma unroll
    for (int l = 0; l < 7; ++l) {
        xo[l] = wrap_index(i - 3 + l, nx);
        yo[l] = wrap_index(j - 3 + l, ny) * nx;
    }
    const long long zb = (long long)xi_stencil_start(k, nz) * plane;
    const long long row = (long long)k * plane + (long long)j * nx;  // this node's x-row
    const long long node = row + i;
    const C* X0 = a.t.xw0 + i * 7;
    const C* X2 = a.t.xw2 + i * 7;
    const C* Y0 = a.t.yw0 + j * 7;
    const C* Y2 = a.t.yw2 + j * 7;
    const C* XI = a.t.xiw + ((long long)k * ny + j) * (NXI * 7);

    auto pop = [&](int q) { return PopReader<ST>{a.fo + (long long)q * a.qs, a.inv_scale[q]}; };

    C F[NQ];
    F[0] = pop(0)(node);
    F[1] = ip_x(pop(1), row, xo, X0);
    F[2] = ip_x(pop(2), row, xo, X2);
    F[3] = ip_xi_y(pop(3), zb + i, yo, plane, Y0, XI + XI3 * 7);
    F[4] = ip_xi_y(pop(4), zb + i, yo, plane, Y2, XI + XI4 * 7);
    F[5] = ip_xi(pop(5), zb + (long long)j * nx + i, plane, XI + XI5 * 7);
    F[6] = ip_xi(pop(6), zb + (long long)j * nx + i, plane, XI + XI6 * 7);
    F[7] = ip_x_y_xi(pop(7), zb, xo, yo, plane, X0, Y0, XI + XI3 * 7);
    F[8] = ip_x_y_xi(pop(8), zb, xo, yo, plane, X2, Y0, XI + XI3 * 7);
    F[9] = ip_x_y_xi(pop(9), zb, xo, yo, plane, X0, Y2, XI + XI4 * 7);
    F[10] = ip_x_y_xi(pop(10), zb, xo, yo, plane, X2, Y2, XI + XI4 * 7);
    F[11] = ip_x_xi(pop(11), zb + (long long)j * nx, xo, plane, X0, XI + XI5 * 7);
    F[12] = ip_x_xi(pop(12), zb + (long long)j * nx, xo, plane, X2, XI + XI5 * 7);
    F[13] = ip_x_xi(pop(13), zb + (long long)j * nx, xo, plane, X0, XI + XI6 * 7);
    F[14] = ip_x_xi(pop(14), zb + (long long)j * nx, xo, plane, X2, XI + XI6 * 7);
    F[15] = ip_y_xi(pop(15), zb + i, yo, plane, Y0, XI + XI15 * 7);
    F[16] = ip_y_xi(pop(16), zb + i, yo, plane, Y2, XI + XI16 * 7);
    F[17] = ip_y_xi(pop(17), zb + i, yo, plane, Y0, XI + XI17 * 7);
    F[18] = ip_y_xi(pop(18), zb + i, yo, plane, Y2, XI + XI18 * 7);

    apply_walls<ST>(F, a, node, i, j, k, xo, yo, plane);

    C tau_eff = (C)a.tau;
    C rho, ux, uy, uz;
    collide_pairs<CD::kDev>(F, (C)*a.rho_modify, (C)a.dt, (C)(*a.force), a.les ? a.t.les_delta[(long long)k * ny + j] : C(0),
                           a.les, (C)a.tau, (C)a.cs, rho, ux, uy, uz, tau_eff, (C)(2.0 * a.tau - 1.0));

    if constexpr (ST == Store::BF16 || ST == Store::FP16) {
        // stochastic rounding dithered by (node, step); thread 0 advances the step counter
        const double ctr = a.step_ctr[a.flux_par];
        const unsigned h = sr_hash((unsigned)node, (unsigned)(unsigned long long)ctr);
#pragma unroll
        for (int q = 0; q < NQ; ++q) a.fn[(long long)q * a.qs + node] = CD::sr(F[q], a.scale[q], h + sr_lane(q));
        if (gid == 0) a.step_ctr[1 - a.flux_par] = ctr + 1.0;
    } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) CD::st(a.fn + (long long)q * a.qs, node, F[q], a.scale[q]);
    }
    a.rho[node] = rho;
    a.u[node] = ux;
    a.v[node] = uy;
    a.w[node] = uz;
}

// ------------------------------------------------------------------------------------
// Bulk flux and forcing controller
// ------------------------------------------------------------------------------------

// Bulk flux through the probe plane in two deterministic passes: one block per z level
// sums its row in a fixed order, then one block sums the levels in index order.
template <class M>
__global__ __launch_bounds__(256) void k_flux_partial(const M* __restrict__ v, const double* __restrict__ fw,
                                                      double* __restrict__ part, int nx, int ny, int jp)
{
    __shared__ double red[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
        s += (double)v[((long long)k * ny + jp) * nx + i] * fw[(long long)k * nx + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[k] = red[0];
}

__global__ __launch_bounds__(256) void k_flux_final(const double* __restrict__ part, int nz, double* acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < nz; k += blockDim.x) s += part[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += red[0];
}

hipError_t launch_flux_probe(hipStream_t st, const void* v, bool fp64, const double* fluxw, double* part, double* acc,
                             int nx, int ny, int nz, int jplane)
{
    if (fp64)
        hipLaunchKernelGGL(k_flux_partial<double>, dim3(nz), dim3(256), 0, st, (const double*)v, fluxw, part, nx, ny,
                           jplane);
    else
        hipLaunchKernelGGL(k_flux_partial<float>, dim3(nz), dim3(256), 0, st, (const float*)v, fluxw, part, nx, ny,
                           jplane);
    hipLaunchKernelGGL(k_flux_final, dim3(1), dim3(256), 0, st, (const double*)part, nz, acc);
    return hipGetLastError();
}

// The probe-plane flux of the current v into slot 0 of the tiled kernel's partials (the
// other slots zero), in a fixed order: each thread sums its levels, then a tree over threads.
template <class M>
__global__ __launch_bounds__(256) void k_flux_refresh(const M* __restrict__ v, const double* __restrict__ dx,
                                                      const double* __restrict__ dz, double* __restrict__ part,
                                                      int nslot, int nx, int ny, int nz, int jp)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long long t = threadIdx.x; t < (long long)nx * nz; t += blockDim.x) {
        const int k = (int)(t / nx), i = (int)(t % nx);
        s += ((double)v[((long long)k * ny + jp) * nx + i] * dx[i]) * dz[k];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nslot; e += blockDim.x) part[e] = e == 0 ? red[0] : 0.0;
}

hipError_t launch_flux_refresh(hipStream_t st, const void* v, bool fp64, const double* dx, const double* dz,
                               double* part, int nslot, int nx, int ny, int nz, int jplane)
{
    if (fp64)
        hipLaunchKernelGGL(k_flux_refresh<double>, dim3(1), dim3(256), 0, st, (const double*)v, dx, dz, part, nslot,
                           nx, ny, nz, jplane);
    else
        hipLaunchKernelGGL(k_flux_refresh<float>, dim3(1), dim3(256), 0, st, (const float*)v, dx, dz, part, nslot,
                           nx, ny, nz, jplane);
    return hipGetLastError();
}

__global__ void k_force_update(double* force, double* acc, double inv_area, int ndtfrc, double beta, double uref)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    // Ub = sum / (LX (LZ - 1)) / NDTFRC;  F += beta (Uref - Ub) Uref / 3.036
    const double ub = acc[0] * inv_area / ndtfrc;
    acc[0] = 0.0;
    force[0] = force[0] + beta * (uref - ub) * uref / 3.036;
}

hipError_t launch_force_update(hipStream_t st, double* force, double* acc, double inv_area, int ndtfrc, double beta,
                               double uref)
{
    hipLaunchKernelGGL(k_force_update, dim3(1), dim3(64), 0, st, force, acc, inv_area, ndtfrc, beta, uref);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Dispatch
// ------------------------------------------------------------------------------------

// ------------------------------------------------------------------------------------
// Two-pass FP64 stream-collide.  The interpolated streaming is separable: every direction
// is a 7-point filter in x and/or y of each level, then a 7-point xi sum over levels with
// the target node's weights.  Pass 1 writes the horizontally filtered values H[q-3] of the
// 16 directions 3-18 at every node (rows at the population stride); pass 2 takes each
// target's xi sums from H (7 levels of one column) and directions 1, 2 into the new buffer;
// pass 3 applies the walls and the collision in place.  The sums keep the
// reference's inner-to-outer order (x, y, then xi) except for 3 and 4, which the reference
// sums xi-first (F3/F4_Intrpl7, interpolationHillISLBM.h:22-58); the difference is FP64
// rounding.
// ------------------------------------------------------------------------------------

// Node of this thread in the two-pass kernels.  A workgroup takes 256 consecutive nodes of
// one level (a chunk of the (i, j) plane).  Workgroups are dealt to the 8 XCDs in turn
// (workgroup b to XCD b % 8), each XCD with its own L2: the workgroups one XCD runs are given
// a contiguous run of the (chunk, level) order, so that the neighbours that share a stencil
// window run on the same L2 — rows of one plane for the x/y filter (chunks fastest), levels
// of one column chunk for the xi sums (levels fastest).
__host__ __device__ inline long long twopass_blocks(int nx, int ny, int nz, int npt = 1)
{
    const long long nch = ((long long)nx * ny + 256 * npt - 1) / (256 * npt);
    return 8 * ((nch * nz + 7) / 8);
}

template <bool kLevelsFastest, int NPT = 1>
__device__ inline bool twopass_node(int nx, int ny, int nz, int& i, int& j, int& k, long long& gid)
{
    const long long plane = (long long)nx * ny;
    const long long nch = (plane + 256 * NPT - 1) / (256 * NPT), nb = nch * nz;
    const long long per = (nb + 7) / 8;
    const long long b = blockIdx.x;
    const long long L = (b % 8) * per + b / 8;
    if (L >= nb) return false;
    const long long c = kLevelsFastest ? L / nz : L % nch;
    k = (int)(kLevelsFastest ? L % nz : L / nch);
    const long long pn = (c * 256 + threadIdx.x) * NPT;
    if (pn >= plane) return false;
    i = (int)(pn % nx);
    j = (int)(pn / nx);
    gid = (long long)k * plane + pn;
    return true;
}

// occupancy targets keep enough waves in flight for the gathers (the compiler would
// otherwise hold every load of the unrolled sums in registers)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_hfilter(StepArgs<Store::F64> a,
                                                                                         double* __restrict__ H)
{
    const int nx = a.nx, ny = a.ny;
    int i, j, k;
    long long gid;
    if (!twopass_node<false>(nx, ny, a.nz, i, j, k, gid)) return;
    const long long plane = (long long)nx * ny;
    int xo[7], yo[7];
#pragma unroll
    for (int l = 0; l < 7; ++l) {
        xo[l] = wrap_index(i - 3 + l, nx);
        yo[l] = wrap_index(j - 3 + l, ny) * nx;
    }
    const long long lev = (long long)k * plane;
    // the x and y weights of the uniform mesh are those of every column and row (to FP64
    // round-off): read once, as wave-uniform values
    using cdouble = const __attribute__((address_space(4))) double;
    const cdouble* X0 = (cdouble*)a.t.xw0;
    const cdouble* X2 = (cdouble*)a.t.xw2;
    const cdouble* Y0 = (cdouble*)a.t.yw0;
    const cdouble* Y2 = (cdouble*)a.t.yw2;
    // one direction at a time: the barrier after each store keeps the next direction's
    // loads from being hoisted, so one 49-tap sum's registers are reused
#pragma unroll
    for (int q = 3; q < NQ; ++q) {
        const double* p = a.fo + (long long)q * a.qs + lev;
        const cdouble* wx = cx(q) > 0 ? X0 : X2;
        const cdouble* wy = cy(q) > 0 ? Y0 : Y2;
        double r;
        if (cx(q) == 0 && cy(q) == 0) {  // 5, 6: at the node
            r = p[(long long)j * nx + i];
        } else if (cy(q) == 0) {  // 11-14: x only
            const double* pr = p + (long long)j * nx;
            r = pr[xo[0]] * wx[0];
#pragma unroll
            for (int l = 1; l < 7; ++l) r = r + pr[xo[l]] * wx[l];
        } else if (cx(q) == 0) {  // 3, 4, 15-18: y only
            const double* pc = p + i;
            r = pc[yo[0]] * wy[0];
#pragma unroll
            for (int n = 1; n < 7; ++n) r = r + pc[yo[n]] * wy[n];
        } else {  // 7-10: x inner, y outer
            r = 0.0;
#pragma unroll
            for (int n = 0; n < 7; ++n) {
                const double* pr = p + yo[n];
                double in = pr[xo[0]] * wx[0];
#pragma unroll
                for (int l = 1; l < 7; ++l) in = in + pr[xo[l]] * wx[l];
                r = (n == 0) ? in * wy[0] : r + in * wy[n];
                asm volatile("" ::: "memory");  // one row of taps in flight at a time
            }
        }
        H[(long long)(q - 3) * a.qs + gid] = r;
        asm volatile("" ::: "memory");
    }
}

// Pass 2: the xi sums of directions 3-18 from H (7 levels of one column) and directions 1, 2;
// the streamed populations go to the new buffer, which the wall and collision pass then
// update in place.  Light in registers, so that many waves keep the gathers in flight.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_xi_stream(StepArgs<Store::F64> a,
                                                                                           const double* __restrict__ H)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    int i, j, k;
    long long gid;
    if (!twopass_node<true>(nx, ny, nz, i, j, k, gid)) return;
    const long long plane = (long long)nx * ny;
    int xo[7], yo[7];
#pragma unroll
    for (int l = 0; l < 7; ++l) {
        xo[l] = wrap_index(i - 3 + l, nx);
        yo[l] = wrap_index(j - 3 + l, ny) * nx;
    }
    const long long row = (long long)k * plane + (long long)j * nx;
    const long long node = row + i;
    const long long col = (long long)xi_stencil_start(k, nz) * plane + (long long)j * nx + i;  // stencil base
    const double* XI = a.t.xiw + ((long long)k * ny + j) * (NXI * 7);
    // xi sum of direction q (H row q-3) with the weights of family fam
    auto xi = [&](int q, int fam) {
        const double* h = H + (long long)(q - 3) * a.qs + col;
        const double* wz = XI + fam * 7;
        double r = h[0] * wz[0];
#pragma unroll
        for (int m = 1; m < 7; ++m) r = r + h[m * plane] * wz[m];
        return r;
    };
    using cdouble = const __attribute__((address_space(4))) double;
    auto ipx = [&](int q, const cdouble* wx) {
        const double* p = a.fo + (long long)q * a.qs + row;
        double r = p[xo[0]] * wx[0];
#pragma unroll
        for (int l = 1; l < 7; ++l) r = r + p[xo[l]] * wx[l];
        return r;
    };
    // groups of four directions, so that the loads in flight stay within the registers
    double F[NQ];
    F[0] = 0.0;  // the collision pass reads direction 0 at the node
    F[1] = ipx(1, (cdouble*)a.t.xw0);
    F[2] = ipx(2, (cdouble*)a.t.xw2);
    asm volatile("" ::: "memory");
    F[3] = xi(3, XI3);
    F[4] = xi(4, XI4);
    F[5] = xi(5, XI5);
    F[6] = xi(6, XI6);
    asm volatile("" ::: "memory");
    F[7] = xi(7, XI3);
    F[8] = xi(8, XI3);
    F[9] = xi(9, XI4);
    F[10] = xi(10, XI4);
    asm volatile("" ::: "memory");
    F[11] = xi(11, XI5);
    F[12] = xi(12, XI5);
    F[13] = xi(13, XI6);
    F[14] = xi(14, XI6);
    asm volatile("" ::: "memory");
    F[15] = xi(15, XI15);
    F[16] = xi(16, XI16);
    F[17] = xi(17, XI17);
    F[18] = xi(18, XI18);
    asm volatile("" ::: "memory");
#pragma unroll
    for (int q = 1; q < NQ; ++q) a.fn[(long long)q * a.qs + node] = F[q];
}

// Pass 3: the walls (bounce-back at k = 0 and nz - 1, BFL at k <= 1; they read only the old
// populations) and the MRT(+LES) collision at each node, in place on the new buffer
// (direction 0 from the old one).
__global__ __launch_bounds__(256) void k_collide_inplace(StepArgs<Store::F64> a)
{
    using CD = Codec<Store::F64>;
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= a.n) return;
    const long long plane = (long long)a.nx * a.ny;
    const int k = (int)(node / plane);
    const int j = (int)((node - (long long)k * plane) / a.nx);
    double F[NQ];
    F[0] = a.fo[node];
#pragma unroll
    for (int q = 1; q < NQ; ++q) F[q] = a.fn[(long long)q * a.qs + node];
    if (k <= 1 || k == a.nz - 1) {
        const int i = (int)(node - (long long)k * plane - (long long)j * a.nx);
        int xo[7], yo[7];
#pragma unroll
        for (int l = 0; l < 7; ++l) {
            xo[l] = wrap_index(i - 3 + l, a.nx);
            yo[l] = wrap_index(j - 3 + l, a.ny) * a.nx;
        }
        apply_walls<Store::F64>(F, a, node, i, j, k, xo, yo, plane);
    }
    double tau_eff = a.tau;
    double rho, ux, uy, uz;
    collide_pairs<CD::kDev>(F, *a.rho_modify, a.dt, *a.force, a.les ? a.t.les_delta[(long long)k * a.ny + j] : 0.0, a.les,
                            a.tau, a.cs, rho, ux, uy, uz, tau_eff, 2.0 * a.tau - 1.0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) CD::st(a.fn + (long long)q * a.qs, node, F[q], a.scale[q]);
    a.rho[node] = rho;
    a.u[node] = ux;
    a.v[node] = uy;
    a.w[node] = uz;
}

// ------------------------------------------------------------------------------------
// Paired forms (even nx): each thread takes two nodes (i, i+1) of one row, i even, and every
// tap is an aligned 16-byte load of a pair.  The x window of the pair is the five pairs
// i-4 .. i+5, which wrap as whole pairs (nx even); node i takes values 1-7 of it, node i+1
// values 2-8.  Same sums, same order as the one-node kernels above.
// ------------------------------------------------------------------------------------
using d2 = HIP_vector_type<double, 2>;

__device__ inline d2 ld2(const double* p) { return *reinterpret_cast<const d2*>(p); }
__device__ inline void st2(double* p, double a, double b) { *reinterpret_cast<d2*>(p) = d2(a, b); }

// LDS form of the xi pass (even nx): a workgroup takes one chunk of 128 nodes (64 pairs) of
// the (i, j) plane at 8 consecutive levels, one wave per level.  For each direction the
// levels the 8 stencils share (at most 14) are staged once into LDS, so each value of H is
// requested from L2 about 1.75 times instead of 7; 28 KB of LDS lets several workgroups share
// a CU and cover each other's barriers.  The rounds are pipelined:
// the next round's values are loaded into registers while the current one is summed.
constexpr int kXiLevels = 8;    // levels per workgroup (one wave each)
constexpr int kXiStage = 14;    // staged levels: 8 + 6
constexpr int kXiDirs = 1;      // directions per staging round (two rounds in LDS at a time)

__host__ __device__ inline long long xi_lds_blocks(int nx, int ny, int nz)
{
    const long long nch = ((long long)nx * ny + 127) / 128;
    const long long ng = (nz + kXiLevels - 1) / kXiLevels;
    return 8 * ((nch * ng + 7) / 8);
}

__global__ __launch_bounds__(64 * kXiLevels) __attribute__((amdgpu_waves_per_eu(4))) void k_xi_stream_lds(StepArgs<Store::F64> a, const double* __restrict__ H)
{
    __shared__ d2 S[2][kXiDirs][kXiStage][64];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const long long plane = (long long)nx * ny;
    const long long nch = (plane + 127) / 128;
    const long long ng = (nz + kXiLevels - 1) / kXiLevels, nb = nch * ng;
    const long long per = (nb + 7) / 8;
    const long long b = blockIdx.x;
    const long long L = (b % 8) * per + b / 8;  // XCD-contiguous run, levels fastest
    if (L >= nb) return;                         // whole workgroup: no barrier skipped
    const long long c = L / ng;
    const int k0 = (int)(L % ng) * kXiLevels;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = k0 + wv;
    const long long pn = c * 128 + 2 * lane;
    const bool active = k < nz && pn < plane;
    const int klast = min(k0 + kXiLevels - 1, nz - 1);
    const int ks0 = xi_stencil_start(k0, nz);
    const int nl = xi_stencil_start(klast, nz) + 7 - ks0;  // staged levels, <= kXiStage
    const int kk = active ? k : klast;
    const int i = (int)(pn % nx), j = (int)(pn / nx);
    const int sl = xi_stencil_start(kk, nz) - ks0;         // first staged level of this stencil
    const double* XI = a.t.xiw + ((long long)kk * ny + (active ? j : 0)) * (NXI * 7);
    const long long row = (long long)kk * plane + (long long)j * nx;
    const long long node = row + i;
    if (active) {  // directions 1 and 2: x filters of the old populations at the node's row
        int xp[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) xp[m] = 2 * wrap_index(i / 2 - 2 + m, nx / 2);
        using cdouble = const __attribute__((address_space(4))) double;
#pragma unroll
        for (int q = 1; q <= 2; ++q) {
            const cdouble* wx = (cdouble*)(q == 1 ? a.t.xw0 : a.t.xw2);
            const double* pr = a.fo + (long long)q * a.qs + row;
            double w[10];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const d2 v = ld2(pr + xp[m]);
                w[2 * m] = v.x;
                w[2 * m + 1] = v.y;
            }
            double r0 = w[1] * wx[0], r1 = w[2] * wx[0];
#pragma unroll
            for (int l = 1; l < 7; ++l) {
                r0 = r0 + w[1 + l] * wx[l];
                r1 = r1 + w[2 + l] * wx[l];
            }
            st2(a.fn + (long long)q * a.qs + node, r0, r1);
        }
    }
    constexpr int fam[NQ] = {0, 0, 0, XI3, XI4, XI5, XI6, XI3, XI3, XI4, XI4, XI5, XI5, XI6, XI6, XI15, XI16, XI17, XI18};
    const long long cbase = (long long)ks0 * plane + c * 128;
    constexpr int kPer = (kXiStage * 64 + 64 * kXiLevels - 1) / (64 * kXiLevels);  // staged pairs per thread
    d2 R[kXiDirs][kPer];
    // levels ks0 .. ks0 + nl - 1 of directions q0 .. q0 + kXiDirs - 1 (pairs past the plane skipped)
    auto fetch = [&](int q0) {
#pragma unroll
        for (int d = 0; d < kXiDirs; ++d) {
            const double* h = H + (long long)(q0 + d - 3) * a.qs + cbase;
#pragma unroll
            for (int it = 0; it < kPer; ++it) {
                const int t = threadIdx.x + it * 64 * kXiLevels, lv = t >> 6, pp = t & 63;
                if (t < nl * 64 && c * 128 + 2 * pp < plane) R[d][it] = ld2(h + (long long)lv * plane + 2 * pp);
            }
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int d = 0; d < kXiDirs; ++d)
#pragma unroll
            for (int it = 0; it < kPer; ++it) {
                const int t = threadIdx.x + it * 64 * kXiLevels;
                if (t < nl * 64) S[buf][d][t >> 6][t & 63] = R[d][it];
            }
    };
    fetch(3);
#pragma unroll
    for (int q0 = 3; q0 < NQ; q0 += kXiDirs) {
        const int buf = ((q0 - 3) / kXiDirs) & 1;
        put(buf);
        // one barrier per round: a wave past it has finished the sums of the round before,
        // whose buffer the next round's put overwrites
        __syncthreads();
        if (q0 + kXiDirs < NQ) fetch(q0 + kXiDirs);
        if (active) {
#pragma unroll
            for (int d = 0; d < kXiDirs; ++d) {
                const double* wz = XI + fam[q0 + d] * 7;
                d2 v = S[buf][d][sl][lane];
                double r0 = v.x * wz[0], r1 = v.y * wz[0];
#pragma unroll
                for (int m = 1; m < 7; ++m) {
                    v = S[buf][d][sl + m][lane];
                    r0 = r0 + v.x * wz[m];
                    r1 = r1 + v.y * wz[m];
                }
                st2(a.fn + (long long)(q0 + d) * a.qs + node, r0, r1);  // streamed populations
            }
        }
    }
}

// LDS form of the x/y filter pass (even nx): a workgroup takes a tile of 64 pairs x 16 rows of
// one level.  For each direction the rows and pairs its taps reach (3 rows and 2 pairs of
// halo, wrapped at load time) are staged once into LDS; 7-10 are filtered in x once per staged
// row, then in y.  The next direction's values are loaded into registers while the current
// one is filtered.  Same sums, same order as the one-node kernel.
constexpr int kHfRows = 16;
constexpr int kHfThreads = 512;  // 8 waves, 2 rows each

__host__ __device__ inline long long hf_lds_blocks(int nx, int ny, int nz)
{
    const long long nb = (long long)((nx / 2 + 63) / 64) * ((ny + kHfRows - 1) / kHfRows) * nz;
    return 8 * ((nb + 7) / 8);
}

__global__ __launch_bounds__(kHfThreads) void k_hfilter_lds(StepArgs<Store::F64> a, double* __restrict__ H)
{
    __shared__ d2 S[kHfRows + 6][68];
    __shared__ d2 X[kHfRows + 6][64];
    const int nx = a.nx, ny = a.ny, nz = a.nz, np = nx / 2;
    const int ns = (np + 63) / 64, nr = (ny + kHfRows - 1) / kHfRows;
    const long long nt = (long long)ns * nr, nb = nt * nz;
    const long long per = (nb + 7) / 8;
    const long long b = blockIdx.x;
    const long long L = (b % 8) * per + b / 8;  // XCD-contiguous run, tiles of one level fastest
    if (L >= nb) return;                         // whole workgroup: no barrier skipped
    const int tile = (int)(L % nt), k = (int)(L / nt);
    const int p0 = (tile % ns) * 64, j0 = (tile / ns) * kHfRows;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int pi = p0 + lane;
    const bool xin = pi < np;
    const long long plane = (long long)nx * ny, lev = (long long)k * plane;
    using cdouble = const __attribute__((address_space(4))) double;
    const cdouble* X0 = (cdouble*)a.t.xw0;
    const cdouble* X2 = (cdouble*)a.t.xw2;
    const cdouble* Y0 = (cdouble*)a.t.yw0;
    const cdouble* Y2 = (cdouble*)a.t.yw2;
    // the staged rows and pairs of direction q: rows r_lo .. r_lo + nrow - 1, pairs c_lo .. c_lo + ncol - 1
    constexpr int kPer = ((kHfRows + 6) * 68 + kHfThreads - 1) / kHfThreads;
    d2 R[kPer];
    auto fetch = [&](int q) {
        const double* p = a.fo + (long long)q * a.qs + lev;
        const bool hx = cx(q) != 0, hy = cy(q) != 0;
        const int r_lo = hy ? 0 : 3, nrow = hy ? kHfRows + 6 : kHfRows;
        const int c_lo = hx ? 0 : 2, ncol = hx ? 68 : 64;
#pragma unroll
        for (int it = 0; it < kPer; ++it) {
            const int t = threadIdx.x + it * kHfThreads;
            if (t < nrow * ncol) {
                const int rr = r_lo + t / ncol, pc = c_lo + t % ncol;
                const int gp = wrap_index(p0 - 2 + pc, np), gj = wrap_index(j0 - 3 + rr, ny);
                R[it] = ld2(p + (long long)gj * nx + 2 * gp);
            }
        }
    };
    auto put = [&](int q) {
        const bool hx = cx(q) != 0, hy = cy(q) != 0;
        const int r_lo = hy ? 0 : 3, nrow = hy ? kHfRows + 6 : kHfRows;
        const int c_lo = hx ? 0 : 2, ncol = hx ? 68 : 64;
#pragma unroll
        for (int it = 0; it < kPer; ++it) {
            const int t = threadIdx.x + it * kHfThreads;
            if (t < nrow * ncol) S[r_lo + t / ncol][c_lo + t % ncol] = R[it];
        }
    };
    fetch(3);
#pragma unroll
    for (int q = 3; q < NQ; ++q) {
        const cdouble* wx = cx(q) > 0 ? X0 : X2;
        const cdouble* wy = cy(q) > 0 ? Y0 : Y2;
        const bool hx = cx(q) != 0, hy = cy(q) != 0;
        put(q);
        __syncthreads();
        if (q + 1 < NQ) fetch(q + 1);  // in flight while this direction is filtered
        if (hx && hy) {  // 7-10: x filter of every staged row
            for (int t = threadIdx.x; t < (kHfRows + 6) * 64; t += kHfThreads) {
                const int rr = t >> 6, pc = t & 63;
                double w[10];
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    const d2 v = S[rr][pc + m];
                    w[2 * m] = v.x;
                    w[2 * m + 1] = v.y;
                }
                double r0 = w[1] * wx[0], r1 = w[2] * wx[0];
#pragma unroll
                for (int l = 1; l < 7; ++l) {
                    r0 = r0 + w[1 + l] * wx[l];
                    r1 = r1 + w[2 + l] * wx[l];
                }
                X[rr][pc] = d2(r0, r1);
            }
            __syncthreads();
        }
#pragma unroll
        for (int rl = 0; rl < kHfRows / (kHfThreads / 64); ++rl) {
            const int r = wv * (kHfRows / (kHfThreads / 64)) + rl, j = j0 + r;
            if (j >= ny || !xin) continue;
            double r0, r1;
            if (!hx && !hy) {  // 5, 6: at the nodes
                const d2 v = S[r + 3][lane + 2];
                r0 = v.x;
                r1 = v.y;
            } else if (!hy) {  // 11-14: x only
                double w[10];
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    const d2 v = S[r + 3][lane + m];
                    w[2 * m] = v.x;
                    w[2 * m + 1] = v.y;
                }
                r0 = w[1] * wx[0];
                r1 = w[2] * wx[0];
#pragma unroll
                for (int l = 1; l < 7; ++l) {
                    r0 = r0 + w[1 + l] * wx[l];
                    r1 = r1 + w[2 + l] * wx[l];
                }
            } else {  // y only (3, 4, 15-18) or y of the x-filtered rows (7-10)
                d2 v = hx ? X[r][lane] : S[r][lane + 2];
                r0 = v.x * wy[0];
                r1 = v.y * wy[0];
#pragma unroll
                for (int n = 1; n < 7; ++n) {
                    v = hx ? X[r + n][lane] : S[r + n][lane + 2];
                    r0 = r0 + v.x * wy[n];
                    r1 = r1 + v.y * wy[n];
                }
            }
            st2(H + (long long)(q - 3) * a.qs + lev + (long long)j * nx + 2 * pi, r0, r1);
        }
        __syncthreads();
    }
}

static hipError_t launch_twopass(hipStream_t st, const StepLaunch& L)
{
    if (L.store != Store::F64 || !L.hbuf) return hipErrorInvalidValue;
    StepArgs<Store::F64> a = make_step_args<Store::F64>(L);
    if (L.g.nx % 2 == 0) {  // paired forms: 16-byte taps
        hipLaunchKernelGGL(k_hfilter_lds, dim3((unsigned)hf_lds_blocks(L.g.nx, L.g.ny, L.g.nz)), dim3(kHfThreads), 0, st, a,
                           (double*)L.hbuf);
        hipLaunchKernelGGL(k_xi_stream_lds, dim3((unsigned)xi_lds_blocks(L.g.nx, L.g.ny, L.g.nz)), dim3(64 * kXiLevels), 0, st, a,
                           (const double*)L.hbuf);
    } else {
        const unsigned nb = (unsigned)twopass_blocks(L.g.nx, L.g.ny, L.g.nz);
        hipLaunchKernelGGL(k_hfilter, dim3(nb), dim3(256), 0, st, a, (double*)L.hbuf);
        hipLaunchKernelGGL(k_xi_stream, dim3(nb), dim3(256), 0, st, a, (const double*)L.hbuf);
    }
    hipLaunchKernelGGL(k_collide_inplace, dim3((unsigned)((L.g.n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <Store ST>
static hipError_t launch_pernode(hipStream_t st, const StepLaunch& L)
{
    StepArgs<ST> a = make_step_args<ST>(L);
    const long long nb = (L.g.n + 255) / 256;
    hipLaunchKernelGGL(k_collide_pernode<ST>, dim3((unsigned)nb), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_tiled(hipStream_t st, const StepLaunch& L);  // islbm_tiled.hip

hipError_t launch_stream_collide(hipStream_t st, const StepLaunch& L)
{
    if (L.kernel == ISLBM_KERNEL_TILED) return launch_tiled(st, L);
    if (L.kernel == ISLBM_KERNEL_TWOPASS) return launch_twopass(st, L);
    switch (L.store) {
    case Store::F64: return launch_pernode<Store::F64>(st, L);
    case Store::F32: return launch_pernode<Store::F32>(st, L);
    case Store::BF16: return launch_pernode<Store::BF16>(st, L);
    case Store::FP16: return launch_pernode<Store::FP16>(st, L);
    }
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------------------------
// Reference padded layout <-> unique nodes (FP64), for the reference-named entry points
// ------------------------------------------------------------------------------------

// The padded layout is [J][K][I] with NX6 = nx + 7, NY6 = ny + 7, NZ6 = nz + 6; padded (I, J, K)
// of unique (i, j, k) is (i + 3, j + 3, k + 3).  Both kernels move one row of one array per
// wave (array blockIdx.y, rows 4 blockIdx.x .. 4 blockIdx.x + 3), so the index arithmetic is
// per wave and 32-bit; a lane moves every 64th value of the row.
__global__ __launch_bounds__(256) void k_padded_to_unique(PtrSet pad, PtrSet uni, int nx, int ny, int nz)
{
    const int row = (int)blockIdx.x * 4 + (int)(threadIdx.x / 64);  // unique row k ny + j
    if (row >= ny * nz) return;
    const int k = row / ny, j = row - k * ny;
    const long long NX6 = nx + 7, NZ6 = nz + 6;
    const double* src = pad.p[blockIdx.y] + (j + 3) * NX6 * NZ6 + (k + 3) * NX6 + 3;
    double* dst = uni.p[blockIdx.y] + (long long)row * nx;
    for (int i = (int)(threadIdx.x % 64); i < nx; i += 64) dst[i] = src[i];
}

// Every padded node of the interior, the periodic duplicates and the three x and y ghost
// layers gets the unique node it is a periodic image of (what periodicSW and the halo
// exchange leave behind, evolution.h:732-777, 980-994); z ghost layers are never read by
// the step and are not written.
__global__ __launch_bounds__(256) void k_unique_to_padded(PtrSet uni, PtrSet pad, int nx, int ny, int nz)
{
    const int NX6 = nx + 7, NY6 = ny + 7;
    const int row = (int)blockIdx.x * 4 + (int)(threadIdx.x / 64);  // padded row k NY6 + J
    if (row >= NY6 * nz) return;
    const int k = row / NY6, J = row - k * NY6;
    const int j = wrap_index(J - 3, ny);
    const double* src = uni.p[blockIdx.y] + ((long long)k * ny + j) * nx;
    double* dst = pad.p[blockIdx.y] + (long long)J * NX6 * (nz + 6) + (long long)(k + 3) * NX6;
    for (int I = (int)(threadIdx.x % 64); I < NX6; I += 64) dst[I] = src[wrap_index(I - 3, nx)];
}

hipError_t launch_padded_to_unique(hipStream_t st, const PtrSet& pad, const PtrSet& uni, int nx, int ny, int nz,
                                   int narr)
{
    const long long rows = (long long)ny * nz;
    if (narr <= 0 || narr > NQ + 4 || rows > 0x7FFFFFFFLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_padded_to_unique, dim3((unsigned)((rows + 3) / 4), (unsigned)narr), dim3(256), 0, st, pad,
                       uni, nx, ny, nz);
    return hipGetLastError();
}

hipError_t launch_unique_to_padded(hipStream_t st, const PtrSet& uni, const PtrSet& pad, int nx, int ny, int nz,
                                   int narr)
{
    const long long rows = (long long)(ny + 7) * nz;
    if (narr <= 0 || narr > NQ + 4 || rows > 0x7FFFFFFFLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_unique_to_padded, dim3((unsigned)((rows + 3) / 4), (unsigned)narr), dim3(256), 0, st, uni,
                       pad, nx, ny, nz);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Seeded equilibrium start on the device (InitialUsingDftFunc, initialization.h:6-49, with
// the seeded velocity perturbation): f = f_eq(rho0, u0 + delta) in FP64, stored in the
// population format; delta_c ~ U(-amp, amp) from the SplitMix64 finaliser of the counter
// ((j nz + k) nx + i) 3 + c.
// ------------------------------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <Store ST>
__global__ __launch_bounds__(256) void k_init_equilibrium(void* f, long long qs, void* mac, long long n, int nx, int ny,
                                                          int nz, double rho0, double u0x, double u0y, double u0z,
                                                          double amp, uint64_t seed, InitScale sc)
{
    using CD = Codec<ST>;
    using raw = typename CD::raw;
    using C = typename CD::C;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((long long)nx * ny));
    const uint64_t base = (((uint64_t)j * nz + k) * nx + i) * 3;
    const double u0[3] = {u0x, u0y, u0z};
    double uu[3];
    for (int c = 0; c < 3; ++c) {
        double d = 0.0;
        if (amp != 0.0) {
            const uint64_t z = splitmix64(seed + (base + c + 1) * 0x9E3779B97F4A7C15ULL);
            d = amp * (2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0);
        }
        uu[c] = u0[c] + d;
    }
    const double udot = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2];
    raw* F = (raw*)f;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double eu = cx(q) * uu[0] + cy(q) * uu[1] + cz(q) * uu[2];
        const double fq = q == 0 ? wq(0) * rho0 * (1.0 - 1.5 * udot)
                                 : wq(q) * rho0 * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * udot);
        if constexpr (ST == Store::F64) {
            F[(long long)q * qs + id] = fq;
        } else if constexpr (ST == Store::F32) {
            F[(long long)q * qs + id] = (float)(fq - wq(q));
        } else {
            // scaled deviation rounded to FP32 first, then to the 16-bit format
            const float v = (float)((fq - wq(q)) * (double)sc.s[q]);
            CD::st(F, (long long)q * qs + id, v, 1.0f);
        }
    }
    C* M = (C*)mac;
    M[id] = (C)rho0;
    M[n + id] = (C)uu[0];
    M[2 * n + id] = (C)uu[1];
    M[3 * n + id] = (C)uu[2];
}

hipError_t launch_init_equilibrium(hipStream_t st, Store store, void* f, long long qs, void* mac, long long n, int nx,
                                   int ny, int nz, double rho0, const double u0[3], double amp, uint64_t seed,
                                   const float* scale)
{
    InitScale sc;
    for (int q = 0; q < NQ; ++q) sc.s[q] = scale ? scale[q] : 1.0f;
    const unsigned nb = (unsigned)((n + 255) / 256);
    switch (store) {
    case Store::F64:
        hipLaunchKernelGGL(k_init_equilibrium<Store::F64>, dim3(nb), dim3(256), 0, st, f, qs, mac, n, nx, ny, nz, rho0, u0[0],
                           u0[1], u0[2], amp, seed, sc);
        break;
    case Store::F32:
        hipLaunchKernelGGL(k_init_equilibrium<Store::F32>, dim3(nb), dim3(256), 0, st, f, qs, mac, n, nx, ny, nz, rho0, u0[0],
                           u0[1], u0[2], amp, seed, sc);
        break;
    case Store::BF16:
        hipLaunchKernelGGL(k_init_equilibrium<Store::BF16>, dim3(nb), dim3(256), 0, st, f, qs, mac, n, nx, ny, nz, rho0,
                           u0[0], u0[1], u0[2], amp, seed, sc);
        break;
    case Store::FP16:
        hipLaunchKernelGGL(k_init_equilibrium<Store::FP16>, dim3(nb), dim3(256), 0, st, f, qs, mac, n, nx, ny, nz, rho0,
                           u0[0], u0[1], u0[2], amp, seed, sc);
        break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Plain streaming copy (float4), the bandwidth reference and counter calibration kernel:
// one float4 per thread, as many workgroups as the copy needs (a grid-stride loop over a
// fixed grid reaches about three quarters of this rate on MI355X)
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst, long long n4)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n4) dst[t] = src[t];
}

hipError_t launch_stream_copy(hipStream_t st, const void* src, void* dst, long long bytes)
{
    const long long n4 = bytes / 16;
    const long long blocks = (n4 + 255) / 256;
    if (blocks <= 0 || blocks > 0x7FFFFFFFLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, st, (const float4*)src, (float4*)dst, n4);
    return hipGetLastError();
}

}  // namespace islbm
