// islbm_step.h -- device building blocks of the stream-collide step: storage codecs,
// 7-point interpolation in the reference's axis orders, wall treatments and the fused MRT
// collision.  Included by the per-node and the tiled kernels.
#pragma once

#include "islbm_internal.h"

#include <type_traits>

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

namespace islbm {

// ------------------------------------------------------------------------------------
// Storage codecs.  kDev: the stored value is the deviation f - w_q (times scale_q).
// ------------------------------------------------------------------------------------
template <Store ST>
struct Codec;

template <>
struct Codec<Store::F64> {
    using raw = double;
    using C = double;
    static constexpr bool kDev = false;
    __device__ static C ld(const raw* p, long long i, float) { return p[i]; }
    __device__ static void st(raw* p, long long i, C v, float) { p[i] = v; }
};

template <>
struct Codec<Store::F32> {
    using raw = float;
    using C = float;
    static constexpr bool kDev = true;
    __device__ static C ld(const raw* p, long long i, float) { return p[i]; }
    __device__ static void st(raw* p, long long i, C v, float) { p[i] = v; }
};

template <>
struct Codec<Store::BF16> {
    using raw = uint16_t;
    using C = float;
    static constexpr bool kDev = true;
    __device__ static C ld(const raw* p, long long i, float isc)
    {
        const uint32_t b = (uint32_t)p[i] << 16;
        return __uint_as_float(b) * isc;
    }
    __device__ static void st(raw* p, long long i, C v, float sc)
    {
        // round to nearest even (v_cvt_pk_bf16_f32)
        p[i] = __builtin_bit_cast(uint16_t, (__bf16)(v * sc));
    }
    // Stochastic rounding (v_cvt_sr_bf16_f32): the magnitude rounds up when its 16 discarded
    // mantissa bits plus the high 16 bits of rnd carry into bit 16, so it rounds up with
    // probability (discarded bits) / 2^16 over uniform rnd, and the stored code is an
    // unbiased estimate of v * sc.  With round to nearest, the rounding error of a smooth
    // field is the same at neighbouring nodes and step after step: it acts as a body force
    // and slows the bulk flow (12% in 4,000 steps at 64 x 128 x 64, Re 10595).
    __device__ static raw sr(C v, float sc, unsigned rnd)
    {
        typedef __attribute__((ext_vector_type(2))) __bf16 bf2;
        const bf2 r = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf2, 0u), v * sc, rnd, false);
        return __builtin_bit_cast(uint16_t, r[0]);
    }
};

// Dither of the stochastic rounding of 16-bit stores (BF16 and FP16): a hash of the unique node index and the
// solve's step count (the lowbias32 integer finaliser), then direction q takes
// hash + q * 0x9E3779B9.  It depends on nothing else, so a solve's stores are the same in
// either kernel form, tile width, alone or grouped, eager or graph-replayed.
__device__ inline unsigned sr_hash(unsigned node, unsigned step)
{
    unsigned x = node * 0x9E3779B1u + step * 0x85EBCA77u;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}
__host__ __device__ constexpr unsigned sr_lane(int q) { return (unsigned)q * 0x9E3779B9u; }

template <>
struct Codec<Store::FP16> {
    using raw = _Float16;
    using C = float;
    static constexpr bool kDev = true;
    __device__ static C ld(const raw* p, long long i, float isc) { return (float)p[i] * isc; }
    __device__ static void st(raw* p, long long i, C v, float sc) { p[i] = (_Float16)(v * sc); }
    // stochastic rounding (v_cvt_sr_f16_f32), as Codec<BF16>::sr
    __device__ static raw sr(C v, float sc, unsigned rnd)
    {
        typedef __attribute__((ext_vector_type(2))) _Float16 h2;
        const h2 r = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(h2, 0u), v * sc, rnd, false);
        return r[0];
    }
};

template <Store ST>
struct StepArgs {
    using CD = Codec<ST>;
    using raw = typename CD::raw;
    using C = typename CD::C;
    const raw* __restrict__ fo;
    raw* __restrict__ fn;
    C* __restrict__ rho;
    C* __restrict__ u;
    C* __restrict__ v;
    C* __restrict__ w;
    DevTables<C> t;
    const double* force;
    const double* rho_modify;
    int nx, ny, nz;
    long long n;
    long long qs;  // population stride (nodes)
    double tau, dt, cs;
    int les;
    float scale[NQ];
    float inv_scale[NQ];
    // bulk-flux probe fused into the tiled kernel (StepLaunch::flux_*)
    double* flux_part;
    const double* flux_dx;
    const double* flux_dz;
    double* flux_acc;
    int flux_plane, flux_par, flux_nslot;
    double* step_ctr;  // 16-bit storage: step counters by buffer parity (StepLaunch::step_ctr), else null
};

template <Store ST>
inline StepArgs<ST> make_step_args(const StepLaunch& L)
{
    using A = StepArgs<ST>;
    using raw = typename A::raw;
    using C = typename A::C;
    A a;
    a.fo = (const raw*)L.fo;
    a.fn = (raw*)L.fn;
    a.rho = (C*)L.rho;
    a.u = (C*)L.u;
    a.v = (C*)L.v;
    a.w = (C*)L.w;
    a.t = *(const DevTables<C>*)L.tables;
    a.force = L.force;
    a.rho_modify = L.rho_modify;
    a.nx = L.g.nx;
    a.ny = L.g.ny;
    a.nz = L.g.nz;
    a.n = L.g.n;
    a.qs = L.g.qs;
    a.tau = L.tau;
    a.dt = L.dt;
    a.cs = L.cs_smag;
    a.les = L.cs_smag > 0.0;
    for (int q = 0; q < NQ; ++q) {
        a.scale[q] = L.scale ? L.scale[q] : 1.0f;
        a.inv_scale[q] = L.scale ? 1.0f / L.scale[q] : 1.0f;
    }
    a.flux_part = L.flux_part;
    a.flux_dx = L.flux_dx;
    a.flux_dz = L.flux_dz;
    a.flux_acc = L.flux_acc;
    a.flux_plane = L.flux_plane;
    a.flux_par = L.flux_par;
    a.flux_nslot = L.flux_nslot;
    a.step_ctr = L.step_ctr;
    return a;
}

// Reads one population array, decoded to compute precision.
template <Store ST>
struct PopReader {
    const typename Codec<ST>::raw* p;
    float isc;
    __device__ typename Codec<ST>::C operator()(long long i) const { return Codec<ST>::ld(p, i, isc); }
};

// ------------------------------------------------------------------------------------
// 7-point Lagrange interpolation (interpolationHillISLBM.h).  Each helper keeps the
// reference's inner-to-outer axis order; Intrpl7 sums left to right.
//   xo[l]: wrapped x index of column i-3+l;  yo[n]: wrapped (j-3+n)*nx;  zb: z stencil base.
// ------------------------------------------------------------------------------------

// F1/F2_Intrpl7 (:14-20): x only, at this node's row
template <class R, class C>
__device__ inline C ip_x(const R& f, long long row, const int* xo, const C* wx)
{
    C r = f(row + xo[0]) * wx[0];
#pragma unroll
    for (int l = 1; l < 7; ++l) r = r + f(row + xo[l]) * wx[l];
    return r;
}

// F5/F6_Intrpl7 (:60-66): xi only
template <class R, class C>
__device__ inline C ip_xi(const R& f, long long base, long long plane, const C* wz)
{
    C r = f(base) * wz[0];
#pragma unroll
    for (int m = 1; m < 7; ++m) r = r + f(base + m * plane) * wz[m];
    return r;
}

// F3/F4_Intrpl7 (:22-58): xi inner, y outer.  base = zb + i.
template <class R, class C>
__device__ inline C ip_xi_y(const R& f, long long base, const int* yo, long long plane, const C* wy, const C* wz)
{
    C r = C(0);
#pragma unroll 1
    for (int n = 0; n < 7; ++n) {
        C in = f(base + yo[n]) * wz[0];
#pragma unroll
        for (int m = 1; m < 7; ++m) in = in + f(base + m * plane + yo[n]) * wz[m];
        r = (n == 0) ? in * wy[0] : r + in * wy[n];
    }
    return r;
}

// Y_XI_Intrpl7 (:150-161): y inner, xi outer.  base = zb + i.
template <class R, class C>
__device__ inline C ip_y_xi(const R& f, long long base, const int* yo, long long plane, const C* wy, const C* wz)
{
    C r = C(0);
#pragma unroll 1
    for (int m = 0; m < 7; ++m) {
        C in = f(base + m * plane + yo[0]) * wy[0];
#pragma unroll
        for (int n = 1; n < 7; ++n) in = in + f(base + m * plane + yo[n]) * wy[n];
        r = (m == 0) ? in * wz[0] : r + in * wz[m];
    }
    return r;
}

// X_XI_Intrpl7 (:137-148): x inner, xi outer.  base = zb + j*nx.
template <class R, class C>
__device__ inline C ip_x_xi(const R& f, long long base, const int* xo, long long plane, const C* wx, const C* wz)
{
    C r = C(0);
#pragma unroll 1
    for (int m = 0; m < 7; ++m) {
        C in = f(base + m * plane + xo[0]) * wx[0];
#pragma unroll
        for (int l = 1; l < 7; ++l) in = in + f(base + m * plane + xo[l]) * wx[l];
        r = (m == 0) ? in * wz[0] : r + in * wz[m];
    }
    return r;
}

// X_Y_XI_Intrpl7 (:68-135): x inner, y middle, xi outer.  base = zb.
template <class R, class C>
__device__ inline C ip_x_y_xi(const R& f, long long base, const int* xo, const int* yo, long long plane, const C* wx,
                              const C* wy, const C* wz)
{
    C r = C(0);
#pragma unroll 1
    for (int m = 0; m < 7; ++m) {
        C mid = C(0);
#pragma unroll
        for (int n = 0; n < 7; ++n) {
            const long long b = base + m * plane + yo[n];
            C in = f(b + xo[0]) * wx[0];
#pragma unroll
            for (int l = 1; l < 7; ++l) in = in + f(b + xo[l]) * wx[l];
            mid = (n == 0) ? in * wy[0] : mid + in * wy[n];
        }
        r = (m == 0) ? mid * wz[0] : r + mid * wz[m];
    }
    return r;
}

// The two BFL gathers of the wall treatment, in the same summation order as ip_y_xi and
// ip_x_y_xi.  They run on the first two levels only, so the xi loop stays rolled: fully
// unrolled, their 49 addresses per xi level would set the register allocation of the whole
// kernel they are inlined into.  Within a xi level the loads are issued together before
// the sums that use them (one memory round trip per level instead of one per tap), and the
// weights come through the constant address space (scalar loads where the row is uniform).
template <class T>
using cptr = const __attribute__((address_space(4))) T*;

template <class R, class C>
__device__ inline C wall_y_xi(const R& f, long long base, const int* yo, long long plane, cptr<C> wy, cptr<C> wz)
{
    C r = C(0);
#pragma unroll 1
    for (int m = 0; m < 7; ++m) {
        C v[7];
#pragma unroll
        for (int n = 0; n < 7; ++n) v[n] = f(base + m * plane + yo[n]);
        C in = v[0] * wy[0];
#pragma unroll
        for (int n = 1; n < 7; ++n) in = in + v[n] * wy[n];
        r = (m == 0) ? in * wz[0] : r + in * wz[m];
    }
    return r;
}

template <class R, class C>
__device__ inline C wall_x_y_xi(const R& f, long long base, const int* xo, const int* yo, long long plane, cptr<C> wx,
                                cptr<C> wy, cptr<C> wz)
{
    C r = C(0);
#pragma unroll 1
    for (int m = 0; m < 7; ++m) {
        C mid = C(0);
#pragma unroll 1
        for (int n = 0; n < 7; ++n) {
            const long long b = base + m * plane + yo[n];
            C v[7];
#pragma unroll
            for (int l = 0; l < 7; ++l) v[l] = f(b + xo[l]);
            C in = v[0] * wx[0];
#pragma unroll
            for (int l = 1; l < 7; ++l) in = in + v[l] * wx[l];
            mid = (n == 0) ? in * wy[0] : mid + in * wy[n];
        }
        r = (m == 0) ? mid * wz[0] : r + mid * wz[m];
    }
    return r;
}

// ------------------------------------------------------------------------------------
// Walls.  Flat-wall half-way bounce-back with the stored (previous post-collision)
// populations of the node (evolution.h:482-496), then BFL on the first two layers above
// the hill (evolution.h:505-572; weights read at layer*ny + j).  bflag holds the branch
// chosen in FP64 on the host: 1 = linear (q > 1/2), 2 = interpolated (q < 1/2),
// 3 = flagged with q == 1/2 (no change), 0 = not flagged.
// In FP32 and 16-bit storage the stored values of the node that any rule of its level
// reads (directions 3-18) are loaded in one batch before the rules run, so a wall level
// costs one memory round trip for them rather than one per branch; the branch tables are
// read through the constant address space.
// ------------------------------------------------------------------------------------
// pop(q): a reader of population q's stored values by node index.  The tiled kernel passes
// one over its buffer resource (32-bit offsets), so the wall code holds no 64-bit address
// per population array in scalar registers.
template <Store ST, class C, class P>
__device__ inline void apply_walls_with(C (&F)[NQ], const StepArgs<ST>& a, const P& pop, long long node, int i, int j,
                                        int k, const int* xo, const int* yo, long long plane)
{
    if (k > 1 && k != a.nz - 1) return;
    auto at = [&](int q) { return pop(q)(node); };
    // FP64 kernels read them where a rule uses them: held together they would raise the
    // register allocation of those kernels
    constexpr bool kBatch = !std::is_same_v<C, double>;
    C o_[NQ];
    if constexpr (kBatch) {
#pragma unroll
        for (int q = 3; q < NQ; ++q) o_[q] = at(q);
    }
    auto o = [&](int q) -> C {
        if constexpr (kBatch)
            return o_[q];
        else
            return at(q);
    };
    if (k == 0) {
        F[5] = o(6); F[11] = o(14); F[12] = o(13); F[15] = o(18); F[16] = o(17);
    }
    if (k == a.nz - 1) {
        F[6] = o(5); F[13] = o(12); F[14] = o(11); F[17] = o(16); F[18] = o(15);
    }
    if (k <= 1) {
        const int nb = 2 * a.ny;
        const int ib = k * a.ny + j;
        const cptr<int> bflag = (cptr<int>)a.t.bflag;
        const cptr<C> bq = (cptr<C>)a.t.bq;
        auto W = [&](int which) { return (cptr<C>)a.t.bw + ((long long)which * nb + ib) * 7; };
        const int f3 = bflag[B3 * nb + ib], f4 = bflag[B4 * nb + ib];
        const int f15 = bflag[B15 * nb + ib], f16 = bflag[B16 * nb + ib];
        if (f3 == 1) {
            const C q = bq[B3 * nb + ib];
            const C c1 = C(1) / (C(2) * q), c2 = (C(2) * q - C(1)) / (C(2) * q);
            F[3] = c1 * o(4) + c2 * o(3);
            F[7] = c1 * o(10) + c2 * o(7);
            F[8] = c1 * o(9) + c2 * o(8);
        } else if (f3 == 2) {
            F[3] = wall_y_xi(pop(4), (long long)i, yo, plane, W(BY378), W(BXI378));
            F[7] = wall_x_y_xi(pop(10), 0LL, xo, yo, plane, W(BX37), W(BY378), W(BXI378));
            F[8] = wall_x_y_xi(pop(9), 0LL, xo, yo, plane, W(BX38), W(BY378), W(BXI378));
        }
        if (f4 == 1) {
            const C q = bq[B4 * nb + ib];
            const C c1 = C(1) / (C(2) * q), c2 = (C(2) * q - C(1)) / (C(2) * q);
            F[4] = c1 * o(3) + c2 * o(4);
            F[9] = c1 * o(8) + c2 * o(9);
            F[10] = c1 * o(7) + c2 * o(10);
        } else if (f4 == 2) {
            F[4] = wall_y_xi(pop(3), (long long)i, yo, plane, W(BY4910), W(BXI4910));
            F[9] = wall_x_y_xi(pop(8), 0LL, xo, yo, plane, W(BX49), W(BY4910), W(BXI4910));
            F[10] = wall_x_y_xi(pop(7), 0LL, xo, yo, plane, W(BX410), W(BY4910), W(BXI4910));
        }
        if (f15 == 1) {
            const C q = bq[B15 * nb + ib];
            F[15] = (C(1) / (C(2) * q)) * o(18) + ((C(2) * q - C(1)) / (C(2) * q)) * o(15);
        } else if (f15 == 2) {
            F[15] = wall_y_xi(pop(18), (long long)i, yo, plane, W(BY15), W(BXI15));
        }
        if (f16 == 1) {
            const C q = bq[B16 * nb + ib];
            F[16] = (C(1) / (C(2) * q)) * o(17) + ((C(2) * q - C(1)) / (C(2) * q)) * o(16);
        } else if (f16 == 2) {
            F[16] = wall_y_xi(pop(17), (long long)i, yo, plane, W(BY16), W(BXI16));
        }
    }
}

template <Store ST, class C>
__device__ inline void apply_walls(C (&F)[NQ], const StepArgs<ST>& a, long long node, int i, int j, int k,
                                   const int* xo, const int* yo, long long plane)
{
    // 16-bit codes: the inverse scale as a literal (store_scale16); 1 for the other formats
    constexpr bool k16 = ST == Store::BF16 || ST == Store::FP16;
    auto pop = [&](int q) {
        return PopReader<ST>{a.fo + (long long)q * a.qs, k16 ? 1.0f / store_scale16(q) : 1.0f};
    };
    apply_walls_with<ST>(F, a, pop, node, i, j, k, xo, yo, plane);
}

// ------------------------------------------------------------------------------------
// Moments, equilibrium and MRT collision with body force (evolution.h:597-632,
// MRT_Process.h:4-179).  With orthogonal M:  f' = f - M^T diag(s_a/|M_a|^2) M (f - f_eq),
// evaluated with the integer rows of M; conserved rows (s = 0) are skipped.
// Body force on the y-moving populations: +-F dt/6 (3, 4), +-F dt/12 (7-10, 15-18).
// kDev: F holds f - w_q and rho is rebuilt as 1 + sum.
// LES: tau* = tau + 3 nu_t/dt, nu_t = (Cs D)^2 |S|, |S| from Q = Pi_neq:Pi_neq in closed form.
// ------------------------------------------------------------------------------------
template <bool kDev, class C>
__host__ __device__ inline void collide_node(C (&F)[NQ], C rho_modify, C dt, C force, C delta, int les, C tau, C cs, C& rho_out,
                                    C& ux_out, C& uy_out, C& uz_out, C& tau_eff)
{
    F[0] = F[0] + rho_modify;
    C s = F[0];
#pragma unroll
    for (int q = 1; q < NQ; ++q) s = s + F[q];
    const C rho = kDev ? C(1) + s : s;
    const C ux = (F[1] + F[7] + F[9] + F[11] + F[13] - (F[2] + F[8] + F[10] + F[12] + F[14])) / rho;
    const C uy = (F[3] + F[7] + F[8] + F[15] + F[17] - (F[4] + F[9] + F[10] + F[16] + F[18])) / rho;
    const C uz = (F[5] + F[11] + F[12] + F[15] + F[16] - (F[6] + F[13] + F[14] + F[17] + F[18])) / rho;
    const C udot = ux * ux + uy * uy + uz * uz;

    C d[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        C eu = C(0);
        if (cx(q)) eu = eu + C(cx(q)) * ux;
        if (cy(q)) eu = eu + C(cy(q)) * uy;
        if (cz(q)) eu = eu + C(cz(q)) * uz;
        const C p = C(3.0) * eu + C(4.5) * eu * eu - C(1.5) * udot;
        const C feq = kDev ? C(wq(q)) * (s + rho * p) : C(wq(q)) * rho * (C(1.0) + p);
        d[q] = F[q] - feq;
    }

    C s_nu = C(1) / tau;
    tau_eff = tau;
    if (les) {
        C pxx = 0, pyy = 0, pzz = 0, pxy = 0, pyz = 0, pxz = 0;
#pragma unroll
        for (int q = 1; q < NQ; ++q) {
            if (cx(q)) pxx += d[q];
            if (cy(q)) pyy += d[q];
            if (cz(q)) pzz += d[q];
            if (cx(q) * cy(q)) pxy += C(cx(q) * cy(q)) * d[q];
            if (cy(q) * cz(q)) pyz += C(cy(q) * cz(q)) * d[q];
            if (cx(q) * cz(q)) pxz += C(cx(q) * cz(q)) * d[q];
        }
        const C Q = pxx * pxx + pyy * pyy + pzz * pzz + C(2) * (pxy * pxy + pyz * pyz + pxz * pxz);
        const C Cd = cs * cs * delta * delta;
        const C b = tau * dt;
        const C S = (sqrt(b * b + C(18.0 * 1.4142135623730951) * Cd * sqrt(Q) / rho) - b) / (C(6) * Cd);
        tau_eff = tau + C(3) * Cd * S / dt;
        s_nu = C(1) / tau_eff;
    }

    C r[NQ];
#pragma unroll
    for (int a = 0; a < NQ; ++a) {
        if (is_conserved(a)) {
            r[a] = C(0);
            continue;
        }
        C acc = C(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int m = mrow(a, q);
            if (m == 1)
                acc = acc + d[q];
            else if (m == -1)
                acc = acc - d[q];
            else if (m != 0)
                acc = acc + C(m) * d[q];
        }
        const C rate = is_stress(a) ? s_nu : C(fixed_rate(a));
        r[a] = acc * (rate / C(mnorm(a)));
    }
    const C fdt = force * dt;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        C acc = C(0);
#pragma unroll
        for (int a = 0; a < NQ; ++a) {
            if (is_conserved(a)) continue;
            const int m = mrow(a, q);
            if (m == 1)
                acc = acc + r[a];
            else if (m == -1)
                acc = acc - r[a];
            else if (m != 0)
                acc = acc + C(m) * r[a];
        }
        C v = F[q] - acc;
        if (cy(q) != 0) {
            const C c = (cx(q) == 0 && cz(q) == 0) ? C(1.0 / 6) : C(1.0 / 12);
            v = (cy(q) > 0) ? v + c * fdt : v - c * fdt;
        }
        F[q] = v;
    }
    rho_out = rho;
    ux_out = ux;
    uy_out = uy;
    uz_out = uz;
}


// ------------------------------------------------------------------------------------
// The same collision evaluated through the D3Q19 opposite-pair structure.  With
// s = f_q + f_qbar and a = f_q - f_qbar for the 9 pairs (1,2) (3,4) (5,6) (7,10) (8,9)
// (11,14) (12,13) (15,18) (16,17), the even moments of (f - f_eq) are sums of s, the odd
// ones sums of a, and the update of a pair is E +- O with E the even-row and O the
// odd-row back-transform.  Algebraically identical to collide_node; about half the
// arithmetic.  The body force enters the odd part of the y-moving pairs.
// ------------------------------------------------------------------------------------
// FP32 on the device uses the hardware reciprocal and square root (about 1 ulp); FP64
// and the host keep the correctly rounded division and square root.
template <class C>
__host__ __device__ inline constexpr bool use_fast_math()
{
#if defined(__HIP_DEVICE_COMPILE__)
    return std::is_same_v<C, float>;
#else
    return false;
#endif
}
template <class C>
__host__ __device__ inline C rcp_(C x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same_v<C, float>) return __builtin_amdgcn_rcpf(x);
#endif
    return C(1) / x;
}
template <class C>
__host__ __device__ inline C sqrt_(C x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same_v<C, float>) return __builtin_amdgcn_sqrtf(x);
#endif
    return sqrt(x);
}

// tau_m = 2 tau - 1, from FP64 (FP32 arithmetic only; see the stress rates below).
template <bool kDev, class C>
__host__ __device__ inline void collide_pairs(C (&F)[NQ], C rho_modify, C dt, C force, C delta, int les, C tau, C cs,
                                              C& rho_out, C& ux_out, C& uy_out, C& uz_out, C& tau_eff, C tau_m)
{
    constexpr bool kFast = use_fast_math<C>();
    F[0] = F[0] + rho_modify;
    C sum = F[0];
#pragma unroll
    for (int q = 1; q < NQ; ++q) sum = sum + F[q];
    const C rho = kDev ? C(1) + sum : sum;
    const C irho = rcp_(rho);
    const C ux = (F[1] + F[7] + F[9] + F[11] + F[13] - (F[2] + F[8] + F[10] + F[12] + F[14])) * irho;
    const C uy = (F[3] + F[7] + F[8] + F[15] + F[17] - (F[4] + F[9] + F[10] + F[16] + F[18])) * irho;
    const C uz = (F[5] + F[11] + F[12] + F[15] + F[16] - (F[6] + F[13] + F[14] + F[17] + F[18])) * irho;
    const C udot = ux * ux + uy * uy + uz * uz;
    // f_eq = w rho (1 + 3 eu + 4.5 eu^2 - 1.5 u^2); deviation form subtracts w
    const C base = kDev ? sum - rho * C(1.5) * udot : rho * (C(1) - C(1.5) * udot);  // rho(1-1.5u^2) [- 1]
    const C r3 = rho * C(3), r45 = rho * C(4.5);
    // pair (q, qbar) with e.u = p:  s_eq = 2 w (base + 4.5 rho p^2),  a_eq = 2 w 3 rho p
    auto pr = [&](int q, int qb, C p, C w2, C& s, C& a) {
        s = F[q] + F[qb] - w2 * (base + r45 * p * p);
        a = F[q] - F[qb] - w2 * (r3 * p);
    };
    const C w18 = C(2.0 / 18.0), w36 = C(2.0 / 36.0);
    C s1, a1, s3, a3, s5, a5, s7, a7, s8, a8, s11, a11, s12, a12, s15, a15, s16, a16;
    pr(1, 2, ux, w18, s1, a1);
    pr(3, 4, uy, w18, s3, a3);
    pr(5, 6, uz, w18, s5, a5);
    pr(7, 10, ux + uy, w36, s7, a7);
    pr(8, 9, uy - ux, w36, s8, a8);
    pr(11, 14, ux + uz, w36, s11, a11);
    pr(12, 13, uz - ux, w36, s12, a12);
    pr(15, 18, uy + uz, w36, s15, a15);
    pr(16, 17, uz - uy, w36, s16, a16);
    const C d0 = F[0] - C(1.0 / 3.0) * base;

    // even moments of f - f_eq
    const C sf = s1 + s3 + s5, se = s7 + s8 + s11 + s12 + s15 + s16;
    const C exy = s7 + s8, exz = s11 + s12, eyz = s15 + s16;
    const C m1 = C(-30) * d0 - C(11) * sf + C(8) * se;
    const C m2 = C(12) * d0 - C(4) * sf + se;
    const C m9 = C(2) * s1 - s3 - s5 + exy + exz - C(2) * eyz;
    const C m10 = C(-4) * s1 + C(2) * (s3 + s5) + exy + exz - C(2) * eyz;
    const C m11 = s3 - s5 + exy - exz;
    const C m12 = C(-2) * s3 + C(2) * s5 + exy - exz;
    const C m13 = s7 - s8, m14 = s15 - s16, m15 = s11 - s12;
    // odd moments
    const C m4 = C(-4) * a1 + a7 - a8 + a11 - a12;
    const C m6 = C(-4) * a3 + a7 + a8 + a15 - a16;
    const C m8 = C(-4) * a5 + a11 + a12 + a15 + a16;
    const C m16 = a7 - a8 - a11 + a12;
    const C m17 = -a7 - a8 + a15 - a16;
    const C m18 = a11 + a12 - a15 - a16;

    C s_nu = rcp_(tau);
    // FP32: the viscosity (tau - 1/2) dt / 3 = (1/s_nu - 1/2) dt / 3 depends on the
    // distance eps = 2 - s_nu of the stress rate from 2, small as tau nears 1/2 (0.004 at
    // tau 0.501).  An FP32 s_nu near 2 is off by up to ~1e-7, a viscosity error of
    // ~1e-7 / eps; so eps itself is carried, from tau_m = 2 tau - 1 formed in FP64:
    // eps = (2 tau - 1) / tau, or with the Smagorinsky increment dl,
    // eps = (2 tau - 1 + 2 dl) / (tau + dl).
    C eps = C(0);
    if constexpr (kFast) eps = tau_m * rcp_(tau);
    tau_eff = tau;
    if (les) {
        const C pxx = s1 + exy + exz, pyy = s3 + exy + eyz, pzz = s5 + exz + eyz;
        const C Q = pxx * pxx + pyy * pyy + pzz * pzz + C(2) * (m13 * m13 + m14 * m14 + m15 * m15);
        const C Cd = cs * cs * delta * delta;
        const C b = tau * dt;
        if constexpr (kFast) {
            // 3 Cd |S| / dt with |S| = (sqrt(b^2 + x) - b) / (6 Cd) in the form without the
            // cancellation of sqrt(b^2 + x) - b (x / b^2 ~ 1e-7 in weak flow, where FP32 would
            // leave |S| mostly rounding noise): x / (2 dt (sqrt(b^2 + x) + b))
            const C x = C(18.0 * 1.4142135623730951) * Cd * sqrt_(Q) * irho;
            const C dl = x * rcp_(C(2) * dt * (sqrt_(b * b + x) + b));
            tau_eff = tau + dl;
            eps = (tau_m + C(2) * dl) * rcp_(tau_eff);
        } else {
            const C S = (sqrt(b * b + C(18.0 * 1.4142135623730951) * Cd * sqrt(Q) * irho) - b) / (C(6) * Cd);
            tau_eff = tau + C(3) * Cd * S / dt;
        }
        s_nu = rcp_(tau_eff);
    }
    // stress rate over the row norm n: s_nu / n, or in FP32 (2 - eps) / n as hi + (lo - eps / n)
    // with hi + lo = 2 / n to FP64 accuracy
    auto srate = [&](C m, auto nc) -> C {
        constexpr double n = decltype(nc)::value;
        if constexpr (kFast) {
            constexpr C hi = C(2.0 / n);
            constexpr C lo = C(2.0 / n - (double)hi);
            return m * hi + m * (lo - eps * C(1.0 / n));
        } else {
            return m * (s_nu * C(1.0 / n));
        }
    };
    using n36 = std::integral_constant<int, 36>;
    using n12 = std::integral_constant<int, 12>;
    using n4 = std::integral_constant<int, 4>;
    // r_a = s_a m_a / |M_a|^2
    const C r1 = m1 * C(1.19 / 2394.0), r2 = m2 * C(1.4 / 252.0);
    const C r4 = m4 * C(1.2 / 40.0), r6 = m6 * C(1.2 / 40.0), r8 = m8 * C(1.2 / 40.0);
    const C r9 = srate(m9, n36{}), r10 = m10 * C(1.4 / 72.0);
    const C r11 = srate(m11, n12{}), r12 = m12 * C(1.4 / 24.0);
    const C r13 = srate(m13, n4{}), r14 = srate(m14, n4{}), r15 = srate(m15, n4{});
    const C r16 = m16 * C(1.5 / 8.0), r17 = m17 * C(1.5 / 8.0), r18 = m18 * C(1.5 / 8.0);

    // back-transform: even part E per pair class, odd part O per pair
    const C fe = C(-11) * r1 - C(4) * r2;   // faces
    const C ee = C(8) * r1 + r2;            // edges
    const C E0 = C(-30) * r1 + C(12) * r2;
    const C E1 = fe + C(2) * r9 - C(4) * r10;
    const C E3 = fe - r9 + C(2) * r10 + r11 - C(2) * r12;
    const C E5 = fe - r9 + C(2) * r10 - r11 + C(2) * r12;
    const C exy_ = ee + r9 + r10 + r11 + r12;
    const C exz_ = ee + r9 + r10 - r11 - r12;
    const C eyz_ = ee - C(2) * r9 - C(2) * r10;
    const C fdt = force * dt;
    const C f6 = fdt * C(1.0 / 6), f12 = fdt * C(1.0 / 12);
    const C O1 = C(-4) * r4;
    const C O3 = C(-4) * r6 - f6;
    const C O5 = C(-4) * r8;
    const C O7 = r4 + r6 + r16 - r17 - f12;
    const C O8 = -r4 + r6 - r16 - r17 - f12;
    const C O11 = r4 + r8 - r16 + r18;
    const C O12 = -r4 + r8 + r16 + r18;
    const C O15 = r6 + r8 + r17 - r18 - f12;
    const C O16 = -r6 + r8 - r17 - r18 + f12;
    F[0] = F[0] - E0;
    auto upd = [&](int q, int qb, C E, C O) {
        F[q] = F[q] - (E + O);
        F[qb] = F[qb] - (E - O);
    };
    upd(1, 2, E1, O1);
    upd(3, 4, E3, O3);
    upd(5, 6, E5, O5);
    upd(7, 10, exy_ + r13, O7);
    upd(8, 9, exy_ - r13, O8);
    upd(11, 14, exz_ + r15, O11);
    upd(12, 13, exz_ - r15, O12);
    upd(15, 18, eyz_ + r14, O15);
    upd(16, 17, eyz_ - r14, O16);
    rho_out = rho;
    ux_out = ux;
    uy_out = uy;
    uz_out = uz;
}

}  // namespace islbm
