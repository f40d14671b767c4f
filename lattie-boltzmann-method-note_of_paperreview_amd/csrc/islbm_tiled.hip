// islbm_tiled.hip -- column-marching, LDS-staged, sum-factorised stream-collide kernel for
// FP32 and 16-bit population storage (FP32 arithmetic, deviation form f - w_q).
//
// Work decomposition.  A workgroup of 256 lanes owns a 64 (x) by 2 (y) tile of node
// columns and marches up a chunk of z levels.  Its four waves form two groups of two
// (one wave per y row in each group); the groups split the 16 directions that need a
// 7-level xi stencil, so a lane keeps a ring of 8 directions x 7 levels in registers:
//     group A: 5 6 7 8 9 10 11 12        group B: 3 4 13 14 15 16 17 18
// For every z level the 14 populations with horizontal neighbours are staged into LDS
// (double-buffered, with a 3-node halo where the direction needs one: 7-10 in x and y,
// 3, 4, 15-18 in y, 11-14 in x; 5 and 6 are read at the lane's own node).  The separable
// x and y Lagrange filters run from LDS (the y pass of 7-10 is shared through LDS) and the
// filtered values enter the ring; once the ring covers a target level's xi stencil, each
// group applies the xi weights of the target node to its 8 directions (the last step of
// the interpolated streaming of interpolationHillISLBM.h).  Targets alternate between the
// groups: the other group hands over its 8 values through LDS, and the owner reads
// directions 0, 1, 2 (no z dependence), applies the walls, collides (collide_pairs,
// evolution.h:597-679) and stores.
//
// FP32 is staged with direct global-to-LDS loads (global_load_lds_dword), one 64-lane row
// (plus a 6-lane halo tail) per load, so no registers hold the prefetch; 16-bit storage
// is staged with global_load_lds_ushort (one zero-extended dword per lane) and decoded in
// place.  Blocks are mapped to tiles so that each XCD (blocks dispatched round-robin over
// 8 XCDs) works on a contiguous slab of tiles, keeping the y-halo rows of neighbouring
// tiles in one L2.  The uniform mesh in x and y makes the x and y weights identical for
// every column and row (to FP64 round-off); the kernel reads them as wave-uniform values.
#include "islbm_internal.h"
#include "islbm_step.h"

namespace islbm {

namespace {

constexpr int TX = 64, TY = 2, NTH = 256;
constexpr int HX = TX + 6, HY = TY + 6;
constexpr int XY_SZ = HY * HX;  // 560: tile of a diagonal direction (x and y halo)
constexpr int Y_SZ = HY * TX;   // 512: y halo only
constexpr int X_SZ = TY * HX;   // 140: x halo only
constexpr int OFF_XY = 0;
constexpr int OFF_Y = OFF_XY + 4 * XY_SZ;
constexpr int OFF_X = OFF_Y + 6 * Y_SZ;
constexpr int STAGE = OFF_X + 4 * X_SZ;  // 5872 staged values per level
constexpr int YP_SZ = TY * HX;           // 140: y pass of a diagonal direction
constexpr int OFF_YP = 2 * STAGE;
constexpr int OFF_XF = OFF_YP + 4 * YP_SZ;       // hand-over of 8 directions x 128 columns
constexpr int LDS_FLOATS = OFF_XF + 8 * TX * TY;  // 13328 floats = 52 KB

__device__ constexpr int kXY[4] = {7, 8, 9, 10};
__device__ constexpr int kY[6] = {3, 4, 15, 16, 17, 18};
__device__ constexpr int kX[4] = {11, 12, 13, 14};
__device__ constexpr int kGA[8] = {5, 6, 7, 8, 9, 10, 11, 12};
__device__ constexpr int kGB[8] = {3, 4, 13, 14, 15, 16, 17, 18};
__host__ __device__ constexpr int xsel(int q) { return cx(q) > 0 ? 0 : 1; }  // 0: X0 (x - dt), 1: X2
__host__ __device__ constexpr int ysel(int q) { return cy(q) > 0 ? 0 : 1; }
// xi family of each interpolated direction (initialization.h:235-242, evolution.h:457-475)
__host__ __device__ constexpr int xifam(int q)
{
    return q == 3 || q == 7 || q == 8 ? XI3 : q == 4 || q == 9 || q == 10 ? XI4 : q == 5 || q == 11 || q == 12 ? XI5
         : q == 6 || q == 13 || q == 14 ? XI6 : q == 15 ? XI15 : q == 16 ? XI16 : q == 17 ? XI17 : XI18;
}

struct TiledGeom {
    int ntx, nty, ntiles, nchunks, kc, nwork, nblocks;
};

TiledGeom tiled_geom(int nx, int ny, int nz)
{
    TiledGeom t;
    t.ntx = (nx + TX - 1) / TX;
    t.nty = (ny + TY - 1) / TY;
    t.ntiles = t.ntx * t.nty;
    // about three resident workgroups per CU on 256 CUs, z chunks of at least 8 levels
    int want = (768 + t.ntiles - 1) / t.ntiles;
    int maxc = nz / 8 > 0 ? nz / 8 : 1;
    t.nchunks = want < 1 ? 1 : (want > maxc ? maxc : want);
    t.kc = (nz + t.nchunks - 1) / t.nchunks;
    t.nchunks = (nz + t.kc - 1) / t.kc;
    t.nwork = t.ntiles * t.nchunks;
    t.nblocks = ((t.nwork + 7) / 8) * 8;
    return t;
}

__device__ inline int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

}  // namespace

template <Store ST>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(3))) void k_collide_tiled(StepArgs<ST> a,
                                                                                                 TiledGeom tg)
{
    using CD = Codec<ST>;
    using raw = typename CD::raw;
    constexpr bool k16 = (ST != Store::F32);
    __shared__ float lds[LDS_FLOATS];

    // XCD-aware work mapping: block b runs on XCD b % 8
    const int per_xcd = tg.nblocks / 8;
    const int wk = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (wk >= tg.nwork) return;
    const int chunk = wk / tg.ntiles;
    const int tile = wk % tg.ntiles;
    const int i0 = (tile % tg.ntx) * TX;
    const int j0 = (tile / tg.ntx) * TY;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const long long plane = (long long)nx * ny;
    const long long N = a.n;
    const int t = threadIdx.x;
    const int lane = t % 64;
    const int wave = __builtin_amdgcn_readfirstlane(t / 64);
    const int grp = wave / TY;  // 0: A, 1: B
    const int ty = wave % TY;
    const int tx = lane;
    const int col = ty * TX + tx;  // column within the tile
    const int i = i0 + tx;
    const int j = j0 + ty;
    const bool active = (i < nx) && (j < ny);
    const int iw = i < nx ? i : i % nx;
    const int jw = j < ny ? j : j % ny;
    const int k0 = chunk * tg.kc;
    const int k1 = min(nz, k0 + tg.kc) - 1;
    const int L0 = xi_stencil_start(k0, nz);
    const int L1 = xi_stencil_start(k1, nz) + 6;
    const unsigned own = (unsigned)(jw * nx + iw);

    // this lane's columns in the staged rows: halo rows start at i0 - 3 (64 lanes + a
    // 6-lane tail), plain rows at i0
    const unsigned xh = (unsigned)wrap_index(i0 - 3 + lane, nx);
    const unsigned xt = (unsigned)wrap_index(i0 - 3 + 64 + (lane < 6 ? lane : 0), nx);
    const unsigned xn = (unsigned)iw;

    // wave-uniform interpolation weights (uniform x and y mesh)
    float wx[2][7], wy[2][7];
#pragma unroll
    for (int l = 0; l < 7; ++l) {
        wx[0][l] = a.t.xw0[l];
        wx[1][l] = a.t.xw2[l];
        wy[0][l] = a.t.yw0[l];
        wy[1][l] = a.t.yw2[l];
    }

    // one staged row: 64 lanes (and the 6-lane tail of a halo row) into LDS
    auto dma = [&](const raw* src, float* ldst) {
        if constexpr (sizeof(raw) == 4)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)ldst, 4, 0, 0);
        else
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)ldst, 2, 0, 0);
    };
    auto dma_row = [&](const raw* rowp, float* ldst, bool halo) {
        dma(rowp + (halo ? xh : xn), ldst);
        if (halo && lane < 6) dma(rowp + xt, ldst + 64);
    };
    // stage level lev into buffer b; rows are split over the four waves
    auto fetch = [&](int lev, int b) {
        const raw* base = a.fo + (long long)lev * plane;
        float* L = lds + b * STAGE;
#pragma unroll
        for (int d = 0; d < 4; ++d) {  // 7-10: HY rows with x halo
            const raw* qb = base + (long long)kXY[d] * N;
            for (int rr = wave; rr < HY; rr += 4)
                dma_row(qb + (long long)wrap1(j0 - 3 + rr, ny) * nx, L + OFF_XY + d * XY_SZ + rr * HX, true);
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) {  // 3, 4, 15-18: HY rows
            const raw* qb = base + (long long)kY[d] * N;
            for (int rr = wave; rr < HY; rr += 4)
                dma_row(qb + (long long)wrap1(j0 - 3 + rr, ny) * nx, L + OFF_Y + d * Y_SZ + rr * TX, false);
        }
        {  // 11-14: TY rows with x halo, one direction per wave
            const int d = wave;
            const raw* qb = base + (long long)(11 + d) * N;
#pragma unroll
            for (int rr = 0; rr < TY; ++rr)
                dma_row(qb + (long long)wrap1(j0 + rr, ny) * nx, L + OFF_X + d * X_SZ + rr * HX, true);
        }
    };
    // 16-bit storage: the DMA wrote raw codes as dwords; decode to FP32 deviations in place
    auto decode = [&](int b) {
        if constexpr (k16) {
            float* L = lds + b * STAGE;
            for (int e = t; e < STAGE; e += NTH) {
                const int q = e < OFF_Y ? kXY[e / XY_SZ] : e < OFF_X ? kY[(e - OFF_Y) / Y_SZ] : kX[(e - OFF_X) / X_SZ];
                const unsigned short code = (unsigned short)(__float_as_uint(L[e]) & 0xFFFFu);
                const raw r = __builtin_bit_cast(raw, code);
                L[e] = CD::ld(&r, 0, a.inv_scale[q]);
            }
        }
    };
    auto ldq = [&](int q, long long lev, unsigned off) -> float {
        return CD::ld(a.fo + (long long)q * N + lev * plane, off, a.inv_scale[q]);
    };

    // 7-level ring of this group's 8 directions (kGA or kGB)
    float g[8][7];
#pragma unroll
    for (int d = 0; d < 8; ++d)
#pragma unroll
        for (int m = 0; m < 7; ++m) g[d][m] = 0.f;

    float z5 = 0.f, z6 = 0.f;
    if (grp == 0) {
        z5 = ldq(5, L0, own);
        z6 = ldq(6, L0, own);
    }
    fetch(L0, 0);
    for (int kk = L0; kk <= L1; ++kk) {
        const int b = (kk - L0) & 1;
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): level kk is in LDS
        __syncthreads();
        const float c5 = z5, c6 = z6;
        const bool more = kk < L1;
        if (more) {
            fetch(kk + 1, b ^ 1);  // in flight while level kk is filtered
            if (grp == 0) {
                z5 = ldq(5, kk + 1, own);
                z6 = ldq(6, kk + 1, own);
            }
        }
        if constexpr (k16) {
            decode(b);
            __syncthreads();
        }
        const float* S = lds + b * STAGE;

        // y pass of 7-10 on the tile's TY rows, all HX columns (shared by the workgroup)
#pragma unroll
        for (int m = 0; m < (4 * YP_SZ + NTH - 1) / NTH; ++m) {
            const int e = t + m * NTH;
            if (e < 4 * YP_SZ) {
                const int d = e / YP_SZ, x = e % YP_SZ;
                const int r = x / HX, c = x % HX;
                const float* src = S + OFF_XY + d * XY_SZ + r * HX + c;
                const bool up = d < 2;  // 7, 8: +y (Y0);  9, 10: -y (Y2)
                float s = src[0] * (up ? wy[0][0] : wy[1][0]);
#pragma unroll
                for (int n = 1; n < 7; ++n) s = s + src[n * HX] * (up ? wy[0][n] : wy[1][n]);
                lds[OFF_YP + e] = s;
            }
        }
        __syncthreads();

        // horizontal filters of this group's directions; shift the ring and append
#pragma unroll
        for (int d = 0; d < 8; ++d)
#pragma unroll
            for (int m = 0; m < 6; ++m) g[d][m] = g[d][m + 1];
        if (grp == 0) {
            g[0][6] = c5;
            g[1][6] = c6;
#pragma unroll
            for (int d = 0; d < 4; ++d) {  // 7-10: x pass over the y-passed rows
                const float* w = wx[xsel(kXY[d])];
                const float* src = lds + OFF_YP + d * YP_SZ + ty * HX + tx;
                float s = src[0] * w[0];
#pragma unroll
                for (int l = 1; l < 7; ++l) s = s + src[l] * w[l];
                g[2 + d][6] = s;
            }
#pragma unroll
            for (int d = 0; d < 2; ++d) {  // 11, 12: x filter
                const float* w = wx[xsel(kX[d])];
                const float* src = S + OFF_X + d * X_SZ + ty * HX + tx;
                float s = src[0] * w[0];
#pragma unroll
                for (int l = 1; l < 7; ++l) s = s + src[l] * w[l];
                g[6 + d][6] = s;
            }
        } else {
#pragma unroll
            for (int d = 0; d < 2; ++d) {  // 3, 4: y filter
                const float* w = wy[ysel(kY[d])];
                const float* src = S + OFF_Y + d * Y_SZ + ty * TX + tx;
                float s = src[0] * w[0];
#pragma unroll
                for (int n = 1; n < 7; ++n) s = s + src[n * TX] * w[n];
                g[d][6] = s;
            }
#pragma unroll
            for (int d = 2; d < 4; ++d) {  // 13, 14: x filter
                const float* w = wx[xsel(kX[d])];
                const float* src = S + OFF_X + d * X_SZ + ty * HX + tx;
                float s = src[0] * w[0];
#pragma unroll
                for (int l = 1; l < 7; ++l) s = s + src[l] * w[l];
                g[d][6] = s;
            }
#pragma unroll
            for (int d = 2; d < 6; ++d) {  // 15-18: y filter
                const float* w = wy[ysel(kY[d])];
                const float* src = S + OFF_Y + d * Y_SZ + ty * TX + tx;
                float s = src[0] * w[0];
#pragma unroll
                for (int n = 1; n < 7; ++n) s = s + src[n * TX] * w[n];
                g[2 + d][6] = s;
            }
        }

        if (kk - L0 < 6) continue;
        // targets whose xi stencil starts at base = kk - 6
        const int base = kk - 6;
        int tlo, thi;
        if (base == 0) {
            tlo = 0; thi = 3;
        } else if (base == nz - 7) {
            tlo = nz - 4; thi = nz - 1;
        } else {
            tlo = thi = base + 3;
        }
        tlo = max(tlo, k0);
        thi = min(thi, k1);
        for (int tk = tlo; tk <= thi; ++tk) {
            const int owner = tk & 1;
            const float* XI = a.t.xiw + ((long long)tk * ny + jw) * (NXI * 7);
            float Fm[8];
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const int q = grp == 0 ? kGA[d] : kGB[d];
                const float* wz = XI + xifam(q) * 7;
                float s = g[d][0] * wz[0];
#pragma unroll
                for (int m = 1; m < 7; ++m) s = s + g[d][m] * wz[m];
                Fm[d] = s;
            }
            if (grp != owner) {
#pragma unroll
                for (int d = 0; d < 8; ++d) lds[OFF_XF + d * (TX * TY) + col] = Fm[d];
            }
            __syncthreads();
            if (grp == owner) {
                float F[NQ];
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const float o = lds[OFF_XF + d * (TX * TY) + col];
                    if (grp == 0) {
                        F[kGA[d]] = Fm[d];
                        F[kGB[d]] = o;
                    } else {
                        F[kGB[d]] = Fm[d];
                        F[kGA[d]] = o;
                    }
                }
                const unsigned row = (unsigned)(jw * nx);
                F[0] = ldq(0, tk, own);
                {
                    float s1 = 0.f, s2 = 0.f;
#pragma unroll
                    for (int l = 0; l < 7; ++l) {
                        const unsigned o = row + (unsigned)wrap_index(iw - 3 + l, nx);
                        s1 = (l == 0) ? ldq(1, tk, o) * wx[0][0] : s1 + ldq(1, tk, o) * wx[0][l];
                        s2 = (l == 0) ? ldq(2, tk, o) * wx[1][0] : s2 + ldq(2, tk, o) * wx[1][l];
                    }
                    F[1] = s1;
                    F[2] = s2;
                }
                const long long node = (long long)tk * plane + own;
                if (tk <= 1 || tk == nz - 1) {
                    int xo[7], yo[7];
#pragma unroll
                    for (int l = 0; l < 7; ++l) {
                        xo[l] = wrap_index(iw - 3 + l, nx);
                        yo[l] = wrap_index(jw - 3 + l, ny) * nx;
                    }
                    apply_walls<ST>(F, a, node, iw, jw, tk, xo, yo, plane);
                }
                float rho, ux, uy, uz, tau_eff;
                collide_pairs<true>(F, (float)*a.rho_modify, (float)a.dt, (float)(*a.force),
                                    a.les ? a.t.les_delta[(long long)tk * ny + jw] : 0.f, a.les, (float)a.tau,
                                    (float)a.cs, rho, ux, uy, uz, tau_eff);
                if (active) {
                    raw* fnb = a.fn + (long long)tk * plane;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) CD::st(fnb + (long long)q * N, own, F[q], a.scale[q]);
                    (a.rho + (long long)tk * plane)[own] = rho;
                    (a.u + (long long)tk * plane)[own] = ux;
                    (a.v + (long long)tk * plane)[own] = uy;
                    (a.w + (long long)tk * plane)[own] = uz;
                }
            }
            if (tk < thi) __syncthreads();  // the hand-over buffer is reused by the next target
        }
    }
}

bool tiled_supported(const LaunchGeom& g, Store store, bool fp64)
{
    return !fp64 && store != Store::F64 && g.nx >= 8 && g.ny >= 8 && g.nz >= 8;
}

template <Store ST>
static hipError_t launch_t(hipStream_t st, const StepLaunch& L)
{
    StepArgs<ST> a = make_step_args<ST>(L);
    TiledGeom tg = tiled_geom(L.g.nx, L.g.ny, L.g.nz);
    hipLaunchKernelGGL(k_collide_tiled<ST>, dim3(tg.nblocks), dim3(NTH), 0, st, a, tg);
    return hipGetLastError();
}

hipError_t launch_tiled(hipStream_t st, const StepLaunch& L)
{
    switch (L.store) {
    case Store::F32: return launch_t<Store::F32>(st, L);
    case Store::BF16: return launch_t<Store::BF16>(st, L);
    case Store::FP16: return launch_t<Store::FP16>(st, L);
    default: return hipErrorNotSupported;
    }
}

}  // namespace islbm
