// islbm_common.h -- lattice constants, hill profile and Lagrange helpers shared by the host
// geometry code and the gfx950 kernels of the ISLBM periodic-hill solver.
//
// Axis convention is the reference's: x = spanwise (LX), y = streamwise (LY),
// z = wall-normal (LZ).  Unique-node arrays are [k][j][i] with i (x) fastest.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace islbm {

constexpr int NQ = 19;

// D3Q19 velocity set: initialization.h:7-9 (0; +-x 1,2; +-y 3,4; +-z 5,6; xy 7-10;
// xz 11-14; yz 15-18).
__host__ __device__ constexpr int cx(int q)
{
    constexpr int v[NQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    return v[q];
}
__host__ __device__ constexpr int cy(int q)
{
    constexpr int v[NQ] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    return v[q];
}
__host__ __device__ constexpr int cz(int q)
{
    constexpr int v[NQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    return v[q];
}
// weights: initialization.h:10-12
__host__ __device__ constexpr double wq(int q) { return q == 0 ? 1.0 / 3.0 : (q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }
// Per-direction scale of the 16-bit storage forms, s_q = 1 / w_q: 3, 18 and 36, exact in
// FP32.  Host and kernels take these constants (kernels as literals, not as arguments held
// in scalar registers).
__host__ __device__ constexpr float store_scale16(int q) { return q == 0 ? 3.0f : q <= 6 ? 18.0f : 36.0f; }
static_assert(store_scale16(0) == 1.0 / wq(0) && store_scale16(3) == 1.0 / wq(3) && store_scale16(9) == 1.0 / wq(9),
              "16-bit scales are 1 / w_q");

// D3Q19 d'Humieres moment basis row a evaluated at velocity q (MRT_Matrix.h:27-46); the
// rows are orthogonal, so M^-1 = M^T diag(|row|^2)^-1 (the MRT_Matrix.h:49-68 table).
__host__ __device__ constexpr int mrow(int a, int q)
{
    const int ex = cx(q), ey = cy(q), ez = cz(q);
    const int c2 = ex * ex + ey * ey + ez * ez;
    switch (a) {
    case 0: return 1;
    case 1: return 19 * c2 - 30;
    case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
    case 3: return ex;
    case 4: return (5 * c2 - 9) * ex;
    case 5: return ey;
    case 6: return (5 * c2 - 9) * ey;
    case 7: return ez;
    case 8: return (5 * c2 - 9) * ez;
    case 9: return 3 * ex * ex - c2;
    case 10: return (3 * c2 - 5) * (3 * ex * ex - c2);
    case 11: return ey * ey - ez * ez;
    case 12: return (3 * c2 - 5) * (ey * ey - ez * ez);
    case 13: return ex * ey;
    case 14: return ey * ez;
    case 15: return ex * ez;
    case 16: return (ey * ey - ez * ez) * ex;
    case 17: return (ez * ez - ex * ex) * ey;
    default: return (ex * ex - ey * ey) * ez;
    }
}
__host__ __device__ constexpr int mnorm(int a)
{
    int s = 0;
    for (int q = 0; q < NQ; ++q) s += mrow(a, q) * mrow(a, q);
    return s;
}
// Relaxation rates, MRT_Matrix.h:4-23; rows 9,11,13,14,15 relax at 1/tau (s_nu).
__host__ __device__ constexpr bool is_conserved(int a) { return a == 0 || a == 3 || a == 5 || a == 7; }
__host__ __device__ constexpr bool is_stress(int a) { return a == 9 || a == 11 || a == 13 || a == 14 || a == 15; }
__host__ __device__ constexpr double fixed_rate(int a)
{
    return a == 1 ? 1.19 : (a == 2 || a == 10 || a == 12) ? 1.4 : (a == 4 || a == 6 || a == 8) ? 1.2 : (a >= 16) ? 1.5 : 0.0;
}

// ERCOFTAC periodic-hill profile H(y) in units of h (model.h:4-68), periodic in y.  alpha
// stretches both flanks; alpha = 1 evaluates the reference expression operation for
// operation (the breakpoints compare y itself, the right flank uses (LY - y)).
__host__ __device__ inline double hill_profile(double Y, double ly, double alpha, int flat)
{
    if (flat) return 0.0;
    double Yb = Y;
    if (Y < 0.0)
        Yb = Y + ly;
    else if (Y > ly)
        Yb = Y - ly;
    const double b9 = (54. / 28.) * (9. / 54.) * alpha, b14 = (54. / 28.) * (14. / 54.) * alpha;
    const double b20 = (54. / 28.) * (20. / 54.) * alpha, b30 = (54. / 28.) * (30. / 54.) * alpha;
    const double b40 = (54. / 28.) * (40. / 54.) * alpha, b54 = (54. / 28.) * (54. / 54.) * alpha;
    // polynomial segment s (0..5) of one flank at abscissa t (in h), coefficients in mm
    auto seg = [](int s, double t) -> double {
        switch (s) {
        case 0: {
            double c = 28. + 0.006775070969851 * t * 28 * t * 28 - 0.0021245277758000 * t * 28 * t * 28 * t * 28;
            return 1. / 28. * (28. < c ? 28. : c);
        }
        case 1: return 1. / 28. * (25.07355893131 + 0.9754803562315 * t * 28 - 0.1016116352781 * t * 28 * t * 28 + 0.001889794677828 * t * 28 * t * 28 * t * 28);
        case 2: return 1. / 28. * (25.79601052357 + 0.8206693007457 * t * 28 - 0.09055370274339 * t * 28 * t * 28 + 0.001626510569859 * t * 28 * t * 28 * t * 28);
        case 3: return 1. / 28. * (40.46435022819 - 1.379581654948 * t * 28 + 0.019458845041284 * t * 28 * t * 28 - 0.0002070318932190 * t * 28 * t * 28 * t * 28);
        case 4: return 1. / 28. * (17.92461334664 + 0.8743920332081 * t * 28 - 0.05567361123058 * t * 28 * t * 28 + 0.0006277731764683 * t * 28 * t * 28 * t * 28);
        default: {
            double c = 56.39011190988 - 2.010520359035 * t * 28 + 0.01644919857549 * t * 28 * t * 28 + 0.00002674976141766 * t * 28 * t * 28 * t * 28;
            return 1. / 28. * (0. > c ? 0. : c);
        }
        }
    };
    const double tl = (alpha == 1.0) ? Yb : Yb / alpha;
    const double tr = (alpha == 1.0) ? (ly - Yb) : (ly - Yb) / alpha;
    double h = 0.0;
    if (Yb <= b9) h = seg(0, tl);
    if (Yb > b9 && Yb <= b14) h = seg(1, tl);
    if (Yb > b14 && Yb <= b20) h = seg(2, tl);
    if (Yb > b20 && Yb <= b30) h = seg(3, tl);
    if (Yb > b30 && Yb <= b40) h = seg(4, tl);
    if (Yb > b40 && Yb <= b54) h = seg(5, tl);
    if (Yb < ly - b40 && Yb >= ly - b54) h = seg(5, tr);
    if (Yb < ly - b30 && Yb >= ly - b40) h = seg(4, tr);
    if (Yb < ly - b20 && Yb >= ly - b30) h = seg(3, tr);
    if (Yb < ly - b14 && Yb >= ly - b20) h = seg(2, tr);
    if (Yb < ly - b9 && Yb >= ly - b14) h = seg(1, tr);
    if (Yb >= ly - b9) h = seg(0, tr);
    return h;
}

// Weight of node p[m] of a 7-node Lagrange interpolant at x; the other nodes enter in
// increasing order (initializationTool.h:60-79).
__host__ __device__ inline double lagrange7_weight(double x, const double* p, int m)
{
    const double pm = p[m];
    double r = 1.0;
    bool first = true;
    for (int n = 0; n < 7; ++n) {
        if (n == m) continue;
        if (first) {
            r = (x - p[n]) / (pm - p[n]);
            first = false;
        } else {
            r = r * (x - p[n]) / (pm - p[n]);
        }
    }
    return r;
}

// Start of the 7-plane xi stencil of node k (unique index): one-sided at both walls
// (evolution.h:421-423, initialization.h:194-200).
__host__ __device__ inline int xi_stencil_start(int k, int nz)
{
    return k <= 3 ? 0 : (k >= nz - 4 ? nz - 7 : k - 3);
}

// xi weights of level k for a departure at xi coordinate pxi, over the 7 nodes of k's
// stencil (xi = all nz node coordinates).  near_wall = 0: the reference's 7-point Lagrange
// weights.  near_wall = 1 (ISLBM_XI_CLAMPED_LINEAR): pxi clamped to [xi[0], xi[nz-1]], and
// on the levels whose stencil is one-sided (k <= 3, k >= nz-4) linear weights on the two
// stencil nodes that bracket pxi.
__host__ __device__ inline int xi_stencil_start(int k, int nz);
__host__ __device__ inline void xi_weights7(double pxi, const double* xi, int k, int nz, int near_wall, double* out7)
{
    const int s = xi_stencil_start(k, nz);
    const double* nodes = xi + s;
    if (near_wall) {
        pxi = pxi < xi[0] ? xi[0] : (pxi > xi[nz - 1] ? xi[nz - 1] : pxi);
        if (k <= 3 || k >= nz - 4) {
            int m = 0;
            while (m < 5 && nodes[m + 1] < pxi) ++m;
            const double t = (pxi - nodes[m]) / (nodes[m + 1] - nodes[m]);
            for (int r = 0; r < 7; ++r) out7[r] = 0.0;
            out7[m] = 1.0 - t;
            out7[m + 1] = t;
            return;
        }
    }
    for (int m = 0; m < 7; ++m) out7[m] = lagrange7_weight(pxi, nodes, m);
}

// FP32 form of one 7-point weight set.  Reduced-precision solves interpolate the deviation
// f - w_q, so the FP32 weights' errors in the moments sum_l w_l (l-3)^p act once per step:
// a sum of 1 + e scales every interpolated deviation by 1 + e (a drift of N e after N
// steps, about 1e-8 per step with rounding to nearest), a first-moment error shifts the
// departure point and a second-moment error changes the interpolation's diffusion.  The
// weights are rounded to nearest, then the FP64 residuals of the sum and of the first two
// moments are moved onto three of the weights (the 3 x 3 Vandermonde solve on their
// offsets, each corrected weight rounded again), keeping the choice of three with the
// smallest remaining |d sum| + |d m1| + |d m2| (about 1e-10 instead of 1e-8); every
// weight stays within a few FP32 ulps of the largest weight of its FP64 value.
__host__ __device__ inline void round_weights7(const double* w, float* out)
{
    float r[7];
    double R0 = 0.0, R1 = 0.0, R2 = 0.0;
    for (int l = 0; l < 7; ++l) {
        r[l] = (float)w[l];
        const double e = w[l] - (double)r[l], x = (double)(l - 3);
        R0 += e;
        R1 += e * x;
        R2 += e * x * x;
    }
    auto cost = [&](const float* v) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int l = 0; l < 7; ++l) {
            const double e = w[l] - (double)v[l], x = (double)(l - 3);
            s0 += e;
            s1 += e * x;
            s2 += e * x * x;
        }
        return fabs(s0) + fabs(s1) + fabs(s2);
    };
    for (int l = 0; l < 7; ++l) out[l] = r[l];
    double best = cost(r);
    for (int a = 0; a < 7; ++a)
        for (int b = a + 1; b < 7; ++b)
            for (int c = b + 1; c < 7; ++c) {
                const double xa = a - 3, xb = b - 3, xc = c - 3;
                const double da = (R2 - (xb + xc) * R1 + xb * xc * R0) / ((xa - xb) * (xa - xc));
                const double db = (R2 - (xa + xc) * R1 + xa * xc * R0) / ((xb - xa) * (xb - xc));
                const double dc = (R2 - (xa + xb) * R1 + xa * xb * R0) / ((xc - xa) * (xc - xb));
                float v[7];
                for (int l = 0; l < 7; ++l) v[l] = r[l];
                v[a] = (float)((double)r[a] + da);
                v[b] = (float)((double)r[b] + db);
                v[c] = (float)((double)r[c] + dc);
                const double cc = cost(v);
                if (cc < best) {
                    best = cc;
                    for (int l = 0; l < 7; ++l) out[l] = v[l];
                }
            }
}

__host__ __device__ inline int wrap_index(int a, int n)
{
    a %= n;
    return a < 0 ? a + n : a;
}

// Departure-point families for the xi weights: (dz, dy) in units of minSize
// (initialization.h:235-242): F3 (0,-1) F4 (0,+1) F5 (-1,0) F6 (+1,0)
// F15 (-1,-1) F16 (-1,+1) F17 (+1,-1) F18 (+1,+1).
enum XiFamily { XI3 = 0, XI4, XI5, XI6, XI15, XI16, XI17, XI18, NXI };
__host__ __device__ constexpr int xi_dz(int f) { return f == XI5 || f == XI15 || f == XI16 ? -1 : (f == XI6 || f == XI17 || f == XI18 ? 1 : 0); }
__host__ __device__ constexpr int xi_dy(int f) { return f == XI3 || f == XI15 || f == XI17 ? -1 : (f == XI4 || f == XI16 || f == XI18 ? 1 : 0); }

// Curved-wall (BFL) tables on the first two node layers above the hill.
enum BflFamily { B3 = 0, B4, B15, B16, NBFL };
enum BflWeights { BX37 = 0, BX38, BY378, BXI378, BX49, BX410, BY4910, BXI4910, BY15, BXI15, BY16, BXI16, NBW };

}  // namespace islbm
