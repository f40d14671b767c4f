// islbm_tiled.hip -- column-marching, LDS-staged, sum-factorised stream-collide kernel for
// FP32 and 16-bit population storage (FP32 arithmetic, deviation form f - w_q).
//
// Work decomposition.  A workgroup owns a 64 (x) by TY (y) tile of node columns (TY = 4 with
// 16-byte staging, 2 with 4-byte staging; grids with 16-byte staging that fill 32-column tiles
// better take 32 x 8 tiles, two rows per wave) and marches up a chunk of z levels.  It has 3 x 4 (or
// 3 x 2) waves in three roles, one wave of each role per y row (per row pair):
//   ring waves, group A and group B: a lane keeps a ring of its group's 8 directions x 7
//       levels in registers
//           group A: 5 6 7 8 9 10 11 12        group B: 3 4 13 14 15 16 17 18
//       (16-bit storage: 5 and 13 swapped, see gq)
//   collision waves: they take the 19 populations of a target from LDS, apply the walls,
//       collide (collide_pairs, evolution.h:597-679) and store; they issue no global load.
// The roles run separate loops with one workgroup barrier per step, so each role's
// registers are allocated for that role alone and the collision of one target overlaps
// the ring work of the next.
//
// Per z level the ring waves stage the 14 populations with horizontal neighbours into LDS
// (three level buffers, loads issued two steps ahead, with a 3-node halo where the
// direction needs one: 7-10 in x and y, 3, 4, 15-18 in y, 11-14 in x; 5 and 6 at the nodes
// only), run the separable x and y Lagrange filters from LDS (group A runs the y pass of
// 7-10 for its own row in a wave-private LDS area) and append the filtered values to the
// ring.  Once the ring covers a target level's xi stencil, each ring wave applies the xi
// weights of the target node to its 8 directions (the last step of the interpolated
// streaming of interpolationHillISLBM.h) and hands the results to the collision waves
// through LDS.  One group also hands over directions 0, 1 and 2 (no z dependence: the node's
// 0 and the x filters of 1 and 2 over the target level's rows, which group B stages): group
// A with FP32 storage, group B with 16-bit storage (whichever balances the groups' work).
//
// Staging uses buffer loads straight into LDS, so no registers hold the prefetch: with
// 16-byte staging one dwordx4 load of 64 lanes fills 64 granules of a level; with 4-byte
// staging one 64-lane row per load, 16-bit storage landing as one zero-extended dword per
// lane, decoded in place by the wave that loaded it.
// Blocks are mapped to tiles so that each XCD (blocks dispatched round-robin over 8 XCDs)
// works on a contiguous slab of tiles, keeping the y-halo rows of neighbouring tiles in
// one L2.  The uniform mesh in x and y makes the x and y weights identical for every
// column and row (to FP64 round-off); the kernel reads them as wave-uniform values.
#include <type_traits>

#include "islbm_internal.h"
#include "islbm_step.h"

namespace islbm {

namespace {

// Tile and LDS layout.  A workgroup owns a TX x TY tile of node columns and has 3 NWR waves:
// ring waves of group A (w = 0 .. NWR-1), of group B (w = NWR .. 2NWR-1) and collision waves
// (w = 2NWR .. 3NWR-1).  A wave covers R = 64 / TX rows of the tile: lane l takes column
// l % TX of row R (w % NWR) + l / TX.  Waves w, w + 4, w + 8 share a SIMD.
//   W = false (4-byte staging, TX = 64, TY = 2): rows with an x halo are 70 wide from
//       x = i0 - 3.
//   W = true (16-byte staging, TY = 4 R): rows with an x halo are TX + 2 GV wide from
//       x = i0 - GV (whole 16-byte granules of GV = 4 FP32 or 8 16-bit values); each SIMD
//       then holds one wave of each role.  A level is
//       staged by NW16 dwordx4 LDS-DMA loads of 64 lanes, lane l of load n filling granule
//       64 n + l (16 bytes at float 4 (64 n + l)).  TX = 64 (one row per wave), or TX = 32
//       (two rows per wave) for grids whose rows fill 32-column tiles but not 64-column ones
//       (nx = 32, 96, ...), where a 64-column tile would leave lanes idle.
// A staged level holds
//   XY: 7-10 [4][HY][HXW]   X: 11-14 [4][TY][HXW]   Z: 5, 6 [2][TY][TX]   Y: 3 4 15-18 [6][HY][TX]
// (16-byte staging: XY, X, Z, Y in that order, contiguous, padded to NW16 granule rows).
template <bool W, int VB = 4, int TXW = 64>  // VB: bytes of a staged value (16-byte staging: 4 FP32, 2 BF16/FP16)
struct Lay {
    static_assert(TXW == 64 || (W && TXW == 32), "32-column tiles: 16-byte staging only");
    static constexpr int TX = TXW;
    static constexpr int HX = TX + 6;
    static constexpr int R = 64 / TX;    // rows per wave
    static constexpr int GV = 16 / VB;  // values per 16-byte granule
    static constexpr int TY = W ? 4 * R : 2;
    static constexpr int NWR = W ? 4 : 2;  // waves per role
    static constexpr int NRW = 2 * NWR;    // ring waves
    static constexpr int NCW = NWR;        // collision waves
    static constexpr int NTH = 64 * (NRW + NCW);
    static constexpr int HY = TY + 6;
    static constexpr int Y_SZ = HY * TX;       // y halo only
    static constexpr int HO_SZ = 19 * TX * TY;  // hand-over: directions 0-18 x TX TY columns
    static constexpr int HXW = W ? TX + 2 * GV : HX;  // x-halo row stride: whole granules from i0 - GV
    static constexpr int XO = W ? GV - 3 : 0;           // column of x = i0 - 3
    static constexpr int XY_SZ = HY * HXW;
    static constexpr int X_SZ = TY * HXW;
    static constexpr int OFF_XY = 0;
    static constexpr int OFF_X = W ? 4 * XY_SZ : 4 * XY_SZ + 6 * Y_SZ;
    static constexpr int OFF_Z = OFF_X + 4 * X_SZ;
    static constexpr int OFF_Y = W ? OFF_Z + 2 * TY * TX : 4 * XY_SZ;
    static constexpr int USED = W ? OFF_Y + 6 * Y_SZ : OFF_Z + 2 * TY * TX;
    static constexpr int NW16 = (USED / GV + 63) / 64;
    static constexpr int STAGE = W ? NW16 * 256 : USED;
    // 16-bit codes with 16-byte staging are decoded once per staged value, by the wave that
    // loaded it, into an FP32 copy of the level (DEC): the filters then read FP32 like the
    // FP32 kernel instead of converting each code at every tap that uses it
    static constexpr bool DEC = W && VB == 2;
    // per-target tables: NBUF slots.  Staged levels in LDS: a level's loads are issued two
    // steps before it is appended; NRAW buffers of staged codes (a decoded level's buffer is
    // free once its wave has decoded it, before the barrier of the step that appends it)
    static constexpr int NBUF = 3;
    static constexpr int NRAW = DEC ? 2 : NBUF;
    static constexpr int DEC_SZ = DEC ? (USED + 3) / 4 * 4 : 0;  // a decoded level, two buffers
    static constexpr int OFF_DEC = NRAW * STAGE;
    static constexpr int OFF_YP = OFF_DEC + 2 * DEC_SZ;  // group A: y pass of 7-10 for its row, 4 x HXW per wave
    static constexpr int OFF_HO = OFF_YP + TY * 4 * HXW;
    // per-target tables staged two steps ahead, NBUF slots (slot = hand-over step % NBUF)
    static constexpr int OFF_XIB = OFF_HO + 2 * HO_SZ;  // xi weights of a target, NBUF x TY rows x 64
    // rows of directions 0, 1, 2 of a target level, NBUF slots x TY rows: 0 at its nodes, 1
    // and 2 with the x halo (staged by group B, read by group A); with 16-byte staging one
    // load stages the R rows of a wave, 64 / R lanes (granules) per row
    static constexpr int P_ROW = W ? (VB == 2 ? 128 : 256 / R) : TX + 2 * HX;
    static constexpr int OFF_P012 = OFF_XIB + NBUF * TY * 64;
    static constexpr int LDS_FLOATS = OFF_P012 + NBUF * TY * P_ROW;
    // workgroups resident at once on 256 CUs: the LDS-DMA staging runs at one per CU
    static constexpr int RESIDENT = W ? 256 : 512;
    static_assert(!W || (OFF_X % GV == 0 && OFF_Z % GV == 0 && OFF_Y % GV == 0 && STAGE % 4 == 0 &&
                         OFF_P012 % 4 == 0 && P_ROW % 4 == 0), "16-byte staging needs granule-aligned regions");
    static_assert(W || TY == 2, "the 4-byte fetch split assumes two rows per tile");
    static_assert(!W || P_ROW >= 4 * (TX / GV + 2 * (HXW / GV)), "rows of 0-2 exceed their slot");
};
// 99 KB (4-byte staging), 157 KB (16-byte staging, FP32), 143 KB (FP32, 32-column tiles),
// 157 KB (BF16, FP16) or 154 KB (BF16, FP16, 32-column tiles): one workgroup per CU
static_assert(Lay<false>::LDS_FLOATS * 4 <= 160 * 1024 && Lay<true>::LDS_FLOATS * 4 <= 160 * 1024 &&
                  Lay<true, 2>::LDS_FLOATS * 4 <= 160 * 1024 && Lay<true, 4, 32>::LDS_FLOATS * 4 <= 160 * 1024 &&
                  Lay<true, 2, 32>::LDS_FLOATS * 4 <= 160 * 1024,
              "LDS budget");
// cache policy of the result stores: non-temporal, streamed past the caches (sc1 nt)
constexpr int kStoreAux = 18;
__host__ __device__ constexpr unsigned vmcnt_imm(int n) { return (n & 15) | ((n >> 4) << 14) | 0x0F70; }

// staged direction of each y-halo LDS region
__host__ __device__ constexpr int qy(int d) { return d < 2 ? 3 + d : 13 + d; }  // 3 4 15 16 17 18
// the 8 directions of ring group G.  SW: directions 5 and 13 swapped between the groups
// (group A: 6 13 7-12, group B: 3 4 5 14 15-18), which balances the groups' per-step work
// with 16-bit storage (group B then also hands over 0-2)
__host__ __device__ constexpr int gq(int G, int d, bool SW)
{
    return G == 0 ? (SW && d == 0 ? 6 : SW && d == 1 ? 13 : 5 + d)
                  : (d < 2 ? 3 + d : SW && d == 2 ? 5 : 11 + d);
}
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

TiledGeom tiled_geom(int nx, int ny, int nz, int tx, int ty, int resident)
{
    TiledGeom t;
    t.ntx = (nx + tx - 1) / tx;
    t.nty = (ny + ty - 1) / ty;
    t.ntiles = t.ntx * t.nty;
    // z chunks of at least 16 levels (each chunk re-reads a 6-level xi halo), as many as
    // minimise the makespan: rounds of `resident` workgroups times the steps of a chunk
    // (the fewest chunks among equal makespans)
    const int maxc = nz / 16 > 0 ? nz / 16 : 1;
    long long best = -1;
    for (int c = 1; c <= maxc; ++c) {
        const int kc = (nz + c - 1) / c;
        const int cc = (nz + kc - 1) / kc;
        const long long rounds = ((long long)t.ntiles * cc + resident - 1) / resident;
        const long long cost = rounds * (kc + 6);
        if (best < 0 || cost < best) {
            best = cost;
            t.kc = kc;
            t.nchunks = cc;
        }
    }
    t.nwork = t.ntiles * t.nchunks;
    t.nblocks = ((t.nwork + 7) / 8) * 8;
    return t;
}

__device__ inline int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Buffer resource over a buffer of populations, a field or a table: 32-bit byte offsets
// from its start.  One resource serves all 19 arrays of a population buffer (the direction
// goes into the scalar offset), so the kernel holds a handful of resources in SGPRs.
__device__ inline __amdgpu_buffer_rsrc_t rsrc_of(const void* base, unsigned nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nbytes, 0x00020000);
}

// One population value through a buffer resource, decoded to / encoded from the FP32
// deviation form.
template <Store ST>
struct BufIO {
    using CD = Codec<ST>;
    using raw = typename CD::raw;
    __device__ static float ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float isc)
    {
        if constexpr (sizeof(raw) == 4) {
            return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
        } else {
            const unsigned short code = __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
            const raw v = __builtin_bit_cast(raw, code);
            return CD::ld(&v, 0, isc);
        }
    }
    // 16-bit storage with stochastic rounding (Codec<BF16>::sr, Codec<FP16>::sr)
    __device__ static void st_sr(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float x, float sc, unsigned rnd)
    {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, CD::sr(x, sc, rnd)), r, (int)voff, (int)soff,
                                              kStoreAux);
    }
    __device__ static void st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float x, float sc)
    {
        raw v;
        CD::st(&v, 0, x, sc);
        if constexpr (sizeof(raw) == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, kStoreAux);
        else
            __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, (int)voff, (int)soff, kStoreAux);
    }
};

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations
// (lgkmcnt(0)) and meets the other waves.  __syncthreads() adds a workgroup release fence,
// which on this target waits for every outstanding global load as well (vmcnt(0)),
// including the next level's LDS prefetch, and would serialise the prefetch with the
// filters.  The kernel shares no global memory between waves, so the fence is not needed;
// the empty asm statements keep the compiler from moving LDS accesses across the barrier.
__device__ inline void lds_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt not waited for
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Step schedule, evaluated identically by every wave of the workgroup.  A step appends at
// most one level to the ring (adv) and hands at most one target level over (ho >= 0);
// targets k0..k1 are handed over in order as soon as the ring holds their xi stencil
// (levels base..base+6, base = xi_stencil_start).  Interior targets take one step each;
// the clamped stencils of the first and last four levels of the column hand over up to
// three more targets from the same ring in steps that append nothing.
struct Sched {
    int top;  // newest level in the ring
    int t;    // next target level
};
__device__ inline bool sched_step(Sched& sc, int k1, int nz, bool& adv, int& ho)
{
    adv = false;
    ho = -1;
    if (sc.t > k1) return false;
    const int need = xi_stencil_start(sc.t, nz) + 6;
    adv = sc.top < need;
    if (adv) ++sc.top;
    if (sc.top == need) ho = sc.t++;
    return true;
}

}  // namespace

// 16-byte buffer load into LDS (gfx950): lane l fills the 16 bytes at ldst + 16 l.  The
// host pass checks the builtin's size operand against host target features, so the call
// exists in the device pass only.
__device__ inline void lds_dma16(__amdgpu_buffer_rsrc_t r, float* ldst, unsigned voff, unsigned soff)
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)ldst, 16, (int)voff, (int)soff, 0,
                                             0);
#endif
}

// Byte offset of staging granule G (16-byte staging) for the tile at (i0, j0): the level's
// offset goes into the load's scalar offset.  Granules past the staged level (padding of
// the last load) repeat granule 0.
template <int VB, int TXW>
__device__ inline unsigned granule_offset(int G, int i0, int j0, int nx, int ny, unsigned nbq)
{
    using LW = Lay<true, VB, TXW>;
    constexpr int TX = LW::TX, TY = LW::TY, HY = LW::HY, Y_SZ = LW::Y_SZ, GV = LW::GV;
    constexpr int RH = LW::HXW / GV, RN = TX / GV;  // granules per row with / without x halo
    auto wx = [&](int x) { return ((x % nx) + nx) % nx; };
    auto wy = [&](int y) { return ((y % ny) + ny) % ny; };
    int q, y, x;
    if (G >= (LW::OFF_Y + 6 * Y_SZ) / GV) G = 0;
    if (G < LW::OFF_X / GV) {  // 7-10: rows j0-3 .. j0+TY+2
        const int row = G / RH, cg = G % RH;
        q = 7 + row / HY;
        y = wy(j0 - 3 + row % HY);
        x = wx(i0 - GV + GV * cg);
    } else if (G < LW::OFF_Z / GV) {  // 11-14: rows j0 .. j0+TY-1
        const int g = G - LW::OFF_X / GV, row = g / RH, cg = g % RH;
        q = 11 + row / TY;
        y = wy(j0 + row % TY);
        x = wx(i0 - GV + GV * cg);
    } else if (G < LW::OFF_Y / GV) {  // 5, 6 at the nodes
        const int g = G - LW::OFF_Z / GV;
        q = 5 + g / (TY * RN);
        y = wy(j0 + (g / RN) % TY);
        x = wx(i0 + GV * (g % RN));
    } else {  // 3 4 15-18: rows j0-3 .. j0+TY+2, no x halo
        const int g = G - LW::OFF_Y / GV, row = g / RN;
        q = qy(row / HY);
        y = wy(j0 - 3 + row % HY);
        x = wx(i0 + GV * (g % RN));
    }
    return (unsigned)q * nbq + (unsigned)VB * (unsigned)(y * nx + x);
}

// Wide addressing (each population array through a resource of its own, from its 64-bit
// base): the 16-byte staging loads of a level, each within the granules of one direction,
// so that a load takes the resource of its array.  The 16 staged directions in region order
// (XY 7-10, X 11-14, Z 5 6, Y 3 4 15-18); load n covers granules [g0, g0 + cnt) of
// direction q, at the same places in LDS as with one resource for all arrays.
template <int VB, int TXW>
struct WideLoads {
    using LW = Lay<true, VB, TXW>;
    __host__ __device__ static constexpr int dir_q(int d)
    {
        return d < 4 ? 7 + d : d < 8 ? 7 + d : d < 10 ? d - 3 : qy(d - 10);
    }
    __host__ __device__ static constexpr int dir_g0(int d)
    {
        return d < 4    ? (LW::OFF_XY + d * LW::XY_SZ) / LW::GV
               : d < 8  ? (LW::OFF_X + (d - 4) * LW::X_SZ) / LW::GV
               : d < 10 ? (LW::OFF_Z + (d - 8) * LW::TY * LW::TX) / LW::GV
                        : (LW::OFF_Y + (d - 10) * LW::Y_SZ) / LW::GV;
    }
    __host__ __device__ static constexpr int dir_n(int d)
    {
        return (d < 4 ? LW::XY_SZ : d < 8 ? LW::X_SZ : d < 10 ? LW::TY * LW::TX : LW::Y_SZ) / LW::GV;
    }
    __host__ __device__ static constexpr int dir_loads(int d) { return (dir_n(d) + 63) / 64; }
    __host__ __device__ static constexpr int count()
    {
        int c = 0;
        for (int d = 0; d < 16; ++d) c += dir_loads(d);
        return c;
    }
    static constexpr int N = count();
    __host__ __device__ static constexpr void of(int n, int& q, int& g0, int& cnt)
    {
        int d = 0;
        while (d < 15 && n >= dir_loads(d)) n -= dir_loads(d++);
        q = dir_q(d);
        g0 = dir_g0(d) + 64 * n;
        cnt = dir_n(d) - 64 * n < 64 ? dir_n(d) - 64 * n : 64;
    }
    // the regions are contiguous and end where the staged level ends
    static_assert(dir_g0(15) + dir_n(15) == LW::USED / LW::GV && dir_g0(4) == dir_g0(3) + dir_n(3) &&
                      dir_g0(8) == dir_g0(7) + dir_n(7) && dir_g0(10) == dir_g0(9) + dir_n(9),
                  "staging regions");
};

// a 16-bit population code as an FP32 value, before its per-direction scale
template <Store ST>
__device__ inline float dec16(unsigned short h)
{
    if constexpr (ST == Store::BF16)
        return __uint_as_float((unsigned)h << 16);
    else
        return (float)__builtin_bit_cast(_Float16, h);
}

// One workgroup's work of the stream-collide step of one solve (bid: its block index
// within that solve's launch).  Inlined into the single-solve kernel and the grouped one.
template <Store ST, bool W, int TXW = 64, bool BIG = false>
__device__ __attribute__((always_inline)) inline void tiled_body(const StepArgs<ST>& a, const TiledGeom& tg,
                                                                  const int bid)
{
    using CD = Codec<ST>;
    using raw = typename CD::raw;
    using IO = BufIO<ST>;
    constexpr unsigned SZ = sizeof(raw);
    using LY = Lay<W, W ? (int)SZ : 4, TXW>;
    constexpr int TX = LY::TX, HX = LY::HX, R = LY::R, NWR = LY::NWR;
    constexpr bool k16 = (ST != Store::F32);
    constexpr bool SW = k16;           // ring group composition (gq)
    constexpr int G012 = k16 ? 1 : 0;  // the group that hands over directions 0-2
    constexpr int HXW = LY::HXW, XO = LY::XO, XY_SZ = LY::XY_SZ, X_SZ = LY::X_SZ;
    constexpr int OFF_XY = LY::OFF_XY, OFF_X = LY::OFF_X, OFF_Y = LY::OFF_Y, OFF_Z = LY::OFF_Z, STAGE = LY::STAGE;
    constexpr int OFF_YP = LY::OFF_YP, OFF_HO = LY::OFF_HO, OFF_XIB = LY::OFF_XIB, OFF_P012 = LY::OFF_P012;
    constexpr int P_ROW = LY::P_ROW, NBUF = LY::NBUF, NRAW = LY::NRAW;
    constexpr bool DEC = LY::DEC;
    constexpr int TY = LY::TY, NRW = LY::NRW, HY = LY::HY, Y_SZ = LY::Y_SZ, HO_SZ = LY::HO_SZ, NW16 = LY::NW16;
    __shared__ __attribute__((aligned(16))) float lds[LY::LDS_FLOATS];
    // per-direction scale of 16-bit codes and its inverse, as literals (1 for FP32 storage)
    auto scl_of = [](int q) { return k16 ? store_scale16(q) : 1.0f; };
    auto isc_of = [](int q) { return k16 ? 1.0f / store_scale16(q) : 1.0f; };

    // XCD-aware work mapping: block b runs on XCD b % 8
    const int per_xcd = tg.nblocks / 8;
    const int wk = (bid % 8) * per_xcd + bid / 8;
    if (wk >= tg.nwork) return;
    const int chunk = wk / tg.ntiles;
    const int tile = wk % tg.ntiles;
    const int i0 = (tile % tg.ntx) * TX;
    const int j0 = (tile / tg.ntx) * TY;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const long long plane = (long long)nx * ny;
    const long long N = a.n;
    const unsigned nbq = SZ * (unsigned)a.qs;  // bytes from one population array to the next; 19 of them < 2^32
    const unsigned xi_bytes = 4u * (unsigned)nz * (unsigned)ny * (NXI * 7);
    const int lane = threadIdx.x % 64;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int k0 = chunk * tg.kc;
    const int k1 = min(nz, k0 + tg.kc) - 1;
    const int L0 = xi_stencil_start(k0, nz);
    const int L1 = xi_stencil_start(k1, nz) + 6;
    // this lane's column and its row within the wave's R rows
    const int tx = R == 1 ? lane : lane % TX;
    const int rl = R == 1 ? 0 : lane / TX;
    const int i = i0 + tx;
    const int iw = i < nx ? i : i % nx;
    // A population buffer is read through one resource for its 19 arrays (the direction's
    // offset goes into the load's offsets), or with wide addressing (BIG: grids whose 19
    // arrays exceed 4 GiB) through one resource per array, from its 64-bit base.
    const unsigned long long nbq64 = (unsigned long long)SZ * (unsigned long long)a.qs;
    const auto rfo = rsrc_of(a.fo, BIG ? nbq : NQ * nbq);
    auto qoff = [&](int q) { return BIG ? 0u : (unsigned)q * nbq; };
    // a wide-addressing resource is built where it is used: the empty asm keeps the compiler
    // from computing the 19 of them once and holding them (76 scalar registers, spilled)
    auto wide_rsrc = [&](const void* base, unsigned long long off, unsigned bytes) {
        const unsigned long long p = (unsigned long long)base + off;
        unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)p), hi = __builtin_amdgcn_readfirstlane((unsigned)(p >> 32));
        asm volatile("" : "+s"(lo), "+s"(hi));
        return rsrc_of((const void*)(((unsigned long long)hi << 32) | lo), bytes);
    };
    auto src = [&](int q) {
        if constexpr (BIG)
            return wide_rsrc(a.fo, (unsigned long long)q * nbq64, nbq);
        else
            return rfo;
    };
    // directions 0, 1, 2 of a target (no z dependence) from its staged rows P: the node's own
    // 0 and the 7 x taps of 1 and 2 (weights w1, w2), in FP32 deviation form
    auto reduce012 = [&](const float* P, const float* w1, const float* w2, float& f0, float& f1, float& f2) {
        const float isc0 = isc_of(0), isc1 = isc_of(1), isc2 = isc_of(2);
        if constexpr (W && k16) {
            // 16-bit codes: direction 0 at the nodes, 1 and 2 with the x halo
            const unsigned short* P16 = (const unsigned short*)P;
            const unsigned short* Q1 = P16 + TX + XO + tx;
            const unsigned short* Q2 = P16 + TX + HXW + XO + tx;
            f0 = dec16<ST>(P16[tx]) * isc0;
            float s1 = dec16<ST>(Q1[0]) * w1[0], s2 = dec16<ST>(Q2[0]) * w2[0];
#pragma unroll
            for (int l = 1; l < 7; ++l) {
                s1 = s1 + dec16<ST>(Q1[l]) * w1[l];
                s2 = s2 + dec16<ST>(Q2[l]) * w2[l];
            }
            f1 = s1 * isc1;
            f2 = s2 * isc2;
            return;
        }
        // 4-byte staging: FP32 as stored, 16-bit codes zero-extended to a dword
        auto dec = [&](float v, float isc) -> float {
            if constexpr (k16) {
                const raw r = __builtin_bit_cast(raw, (unsigned short)(__float_as_uint(v) & 0xFFFFu));
                return CD::ld(&r, 0, isc);
            } else {
                return v;
            }
        };
        f0 = dec(P[tx], isc0);
        const float* P1 = P + TX + XO + tx;
        const float* P2 = P + TX + HXW + XO + tx;
        float s1 = dec(P1[0], isc1) * w1[0], s2 = dec(P2[0], isc2) * w2[0];
#pragma unroll
        for (int l = 1; l < 7; ++l) {
            s1 = s1 + dec(P1[l], isc1) * w1[l];
            s2 = s2 + dec(P2[l], isc2) * w2[l];
        }
        f1 = s1;
        f2 = s2;
    };

    if (wave < NRW) {
        // ------------------------------------------------------------------ ring waves
        const int w = wave;  // fetch share
        const int wr = w % NWR;
        const int ty = R * wr + rl;  // row of this lane
        const int col = ty * TX + tx;
        const int j = j0 + ty;
        const int jw = j < ny ? j : j % ny;
        // Staging loads are buffer loads into LDS: a resource per population array, the
        // row offset in an SGPR and this lane's column offset in a VGPR that stays live for
        // the whole kernel (per-load 64-bit VGPR addresses would be recycled by the
        // compiler right after the load, and a recycled address register of an LDS load
        // makes it wait for that load).  A halo row (70 columns from i0 - 3) is two 64-lane
        // loads, columns [i0-3, i0+61) and [i0+3, i0+67) placed at row offsets 0 and 6: the
        // overlap is written twice with the same values.
        const unsigned xh = SZ * (unsigned)wrap_index(i0 - 3 + lane, nx);
        const unsigned xh6 = SZ * (unsigned)wrap_index(i0 + 3 + lane, nx);
        const unsigned xn = SZ * (unsigned)iw;
        unsigned yrow[HY];  // byte offset of staged row rr in a plane
#pragma unroll
        for (int rr = 0; rr < HY; ++rr) yrow[rr] = SZ * (unsigned)(wrap1(j0 - 3 + rr, ny) * nx);
        // 16-byte staging: this wave issues loads n = w, w + NRW, ... (< NL) of each level.  With
        // wide addressing load n covers granules [wg0, wg0 + wcnt) of direction wq (wave-uniform)
        // and the lane's offset is within that direction's array.
        using WL = WideLoads<W ? (int)SZ : 4, TX>;
        constexpr int NL = BIG && W ? WL::N : NW16;
        unsigned g16[(NL + NRW - 1) / NRW];
        int wq[(NL + NRW - 1) / NRW], wg0[(NL + NRW - 1) / NRW], wcnt[(NL + NRW - 1) / NRW];
        if constexpr (W) {
#pragma unroll
            for (int m = 0; m < (NL + NRW - 1) / NRW; ++m) {
                if constexpr (BIG) {
                    int q_, g0_, cnt_;
                    WL::of(min(w + NRW * m, NL - 1), q_, g0_, cnt_);
                    wq[m] = __builtin_amdgcn_readfirstlane(q_);
                    wg0[m] = __builtin_amdgcn_readfirstlane(g0_);
                    wcnt[m] = __builtin_amdgcn_readfirstlane(cnt_);
                    g16[m] = granule_offset<W ? (int)SZ : 4, TX>(wg0[m] + lane, i0, j0, nx, ny, 0u);
                } else {
                    g16[m] = granule_offset<W ? (int)SZ : 4, TX>(64 * (w + NRW * m) + lane, i0, j0, nx, ny, nbq);
                }
            }
        }
        float wx[2][7], wy[2][7];
#pragma unroll
        for (int l = 0; l < 7; ++l) {
            wx[0][l] = a.t.xw0[l];
            wx[1][l] = a.t.xw2[l];
            wy[0][l] = a.t.yw0[l];
            wy[1][l] = a.t.yw2[l];
        }

        auto dma = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float* ldst) {
            auto l = (__attribute__((address_space(3))) void*)ldst;
            if constexpr (SZ == 4)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 4, (int)voff, (int)soff, 0, 0);
            else
                __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 2, (int)voff, (int)soff, 0, 0);
        };
        // this wave's share of level lev: direction 7+w (8 halo rows), 11+w (2 halo rows)
        // and rows 2w, 2w+1 of each y-halo direction
        auto fetch = [&](int lev, int b) {
            const unsigned lv = SZ * (unsigned)lev * (unsigned)plane;
            float* L = lds + b * STAGE;  // b < NRAW
            if constexpr (W && BIG) {
#pragma unroll
                for (int m = 0; m < (NL + NRW - 1) / NRW; ++m) {
                    if (w + NRW * m < NL)
                        if (lane < wcnt[m])  // a direction's last load leaves lanes idle
                            lds_dma16(src(wq[m]), L + 4 * wg0[m], g16[m], lv);
                }
                return;
            } else if constexpr (W) {
#pragma unroll
                for (int m = 0; m < (NW16 + NRW - 1) / NRW; ++m) {
                    const int n = w + NRW * m;
                    if (n < NW16)
                        if (64 * n + lane < LY::USED / LY::GV)  // the last load's padding lanes stay idle
                            lds_dma16(rfo, L + 256 * n, g16[m], lv);
                }
                return;
            }
            const unsigned od = qoff(7 + w) + lv;
#pragma unroll
            for (int rr = 0; rr < HY; ++rr) {
                float* dst = L + OFF_XY + w * XY_SZ + rr * HX;
                dma(src(7 + w), xh, od + yrow[rr], dst);
                dma(src(7 + w), xh6, od + yrow[rr], dst + 6);
            }
            const unsigned ox = qoff(11 + w) + lv;
#pragma unroll
            for (int rr = 0; rr < TY; ++rr) {
                float* dst = L + OFF_X + w * X_SZ + rr * HX;
                dma(src(11 + w), xh, ox + yrow[3 + rr], dst);
                dma(src(11 + w), xh6, ox + yrow[3 + rr], dst + 6);
            }
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                const unsigned oy = qoff(qy(d)) + lv;
#pragma unroll
                for (int r = 0; r < 2; ++r)
                    dma(src(qy(d)), xn, oy + yrow[2 * w + r], L + OFF_Y + d * Y_SZ + (2 * w + r) * TX);
            }
            if (w >= TY) {  // group B: 5 and 6 at the nodes of row w - TY, for group A's ring
                dma(src(5), xn, qoff(5) + lv + yrow[3 + w - TY], L + OFF_Z + (w - TY) * TX);
                dma(src(6), xn, qoff(6) + lv + yrow[3 + w - TY], L + OFF_Z + (TY + w - TY) * TX);
            }
        };
        // 16-bit storage: decode this wave's share to FP32 deviations in place
        auto decode_region = [&](float* R, int n, float isc) {
            if constexpr (k16 && !W) {
                for (int e = lane; e < n; e += 64) {
                    const unsigned short code = (unsigned short)(__float_as_uint(R[e]) & 0xFFFFu);
                    const raw r = __builtin_bit_cast(raw, code);
                    R[e] = CD::ld(&r, 0, isc);
                }
            }
        };
        // 16-byte staging of 16-bit codes: this wave's granules of raw buffer b, 8 codes each,
        // decoded (before their direction's scale) into decoded buffer db
        auto decode16 = [&](int b, int db) {
            if constexpr (DEC) {
                const float* R = lds + b * STAGE;
                float* D = lds + LY::OFF_DEC + db * LY::DEC_SZ;
#pragma unroll
                for (int m = 0; m < (NL + NRW - 1) / NRW; ++m) {
                    const int n = w + NRW * m;
                    const int G = BIG ? wg0[m] + lane : 64 * n + lane;
                    if (n < NL && (BIG ? lane < wcnt[m] : G < LY::USED / LY::GV)) {
                        const uint4 c = *(const uint4*)(R + 4 * G);
                        const unsigned cw[4] = {c.x, c.y, c.z, c.w};
                        float v[8];
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            v[2 * e] = dec16<ST>((unsigned short)(cw[e] & 0xFFFFu));
                            v[2 * e + 1] = dec16<ST>((unsigned short)(cw[e] >> 16));
                        }
                        *(float4*)(D + 8 * G) = make_float4(v[0], v[1], v[2], v[3]);
                        *(float4*)(D + 8 * G + 4) = make_float4(v[4], v[5], v[6], v[7]);
                    }
                }
            }
        };
        auto decode = [&](int b) {
            if constexpr (k16 && !W) {
                float* L = lds + b * STAGE;
                decode_region(L + OFF_XY + w * XY_SZ, XY_SZ, isc_of(7 + w));
                decode_region(L + OFF_X + w * X_SZ, X_SZ, isc_of(11 + w));
#pragma unroll
                for (int d = 0; d < 6; ++d)
                    decode_region(L + OFF_Y + d * Y_SZ + 2 * w * TX, 2 * TX, isc_of(qy(d)));
                if (w >= TY) {
                    decode_region(L + OFF_Z + (w - TY) * TX, TX, isc_of(5));
                    decode_region(L + OFF_Z + (TY + w - TY) * TX, TX, isc_of(6));
                }
            }
        };

        auto ring = [&](auto gc) {
            constexpr int G = decltype(gc)::value;
            // The ring of direction d holds its last 7 appended levels in 8 slots, four
            // register pairs h[d][0..3].  Appends alternate in parity: an even append writes
            // slot 7 (window: slots 1-7); an odd append moves the pairs down by one (three
            // 64-bit moves) and writes slot 6 (window: slots 0-6).  The step body is compiled
            // for both parities, so no append shifts the ring value by value.
            float h[8][8];
#pragma unroll
            for (int d = 0; d < 8; ++d)
#pragma unroll
                for (int m = 0; m < 8; ++m) h[d][m] = 0.f;
            // Every vector-memory instruction of a ring wave is a staging load into LDS, issued
            // two steps before it is read.  In step s, group B stages the xi weights and the
            // rows of 0-2 of the target handed over in step s+2, then every ring wave issues
            // its share of the level two above the one it appends (at the top of the column,
            // level L1 again into the free buffer).  Loads complete in order, so a step waits
            // until only the loads of the step before are outstanding (vmcnt = their count).
            const int nbatch = W ? (NL - w + NRW - 1) / NRW : (G == 0 ? 32 : 34);
            constexpr int nP012 = W ? (BIG ? 3 : 1) : 5;
            constexpr int ntab = G == 1 ? R + nP012 : 0;  // xi of the wave's rows + rows of 0-2, every step
            auto wait_staged = [&](int outstanding) {
                // the counts a ring wave can leave outstanding (0 waits for everything)
                switch (outstanding) {
                case 2: __builtin_amdgcn_s_waitcnt(vmcnt_imm(2)); break;
                case 3: __builtin_amdgcn_s_waitcnt(vmcnt_imm(3)); break;
                case 4: __builtin_amdgcn_s_waitcnt(vmcnt_imm(4)); break;
                case 5: __builtin_amdgcn_s_waitcnt(vmcnt_imm(5)); break;
                case 6: __builtin_amdgcn_s_waitcnt(vmcnt_imm(6)); break;
                case 7: __builtin_amdgcn_s_waitcnt(vmcnt_imm(7)); break;
                case 8: __builtin_amdgcn_s_waitcnt(vmcnt_imm(8)); break;
                case 9: __builtin_amdgcn_s_waitcnt(vmcnt_imm(9)); break;
                case 10: __builtin_amdgcn_s_waitcnt(vmcnt_imm(10)); break;
                case 32: __builtin_amdgcn_s_waitcnt(vmcnt_imm(32)); break;
                case 40: __builtin_amdgcn_s_waitcnt(vmcnt_imm(40)); break;
                default: __builtin_amdgcn_s_waitcnt(vmcnt_imm(0)); break;
                }
            };
            fetch(L0, 0);
            int b = 0;   // stage buffer of the next level to append (level L0 + a in buffer a % NRAW)
            int db = 0;  // its decoded buffer (16-bit codes, 16-byte staging): level L0 + a in a % 2
            int s = 0;  // step
            Sched sc{L0 - 1, k0};
            bool adv;
            int ho;
            // xi weights of each step's target are staged into LDS two steps ahead by the
            // group B waves (one load of the 56 weights of its row by lanes 0-55)
            const auto rxi = rsrc_of(a.t.xiw, xi_bytes);
            unsigned p16 = 0;  // 16-byte staging of directions 0-2 (group B): this lane's granule
            unsigned p16q[3] = {0, 0, 0};  // the same with wide addressing: per direction, within its array
            if constexpr (G == 1 && W && BIG) {
                constexpr int GV = LY::GV;
                const int xs[3] = {i0 + GV * tx, i0 - GV + GV * tx, i0 - GV + GV * tx};
#pragma unroll
                for (int q = 0; q < 3; ++q) p16q[q] = SZ * (unsigned)(jw * nx + ((xs[q] % nx) + nx) % nx);
            } else if constexpr (G == 1 && W) {
                // granules: direction 0 at the nodes, then 1 and 2 with the x halo from i0 - GV;
                // lanes TX rl .. TX rl + TX - 1 stage row rl of the wave
                constexpr int GV = LY::GV, n0 = TX / GV, n12 = HXW / GV;
                const int g = tx < n0 + 2 * n12 ? tx : 0;
                const int q = g < n0 ? 0 : (g < n0 + n12 ? 1 : 2);
                const int x = g < n0 ? i0 + GV * g : i0 - GV + GV * (g - (q == 1 ? n0 : n0 + n12));
                p16 = (unsigned)q * nbq + SZ * (unsigned)(jw * nx + ((x % nx) + nx) % nx);
            }
            auto fetch_xi = [&](int tk, int par) {
                if constexpr (G == 1) {  // an FP32 table whatever the population storage
#pragma unroll
                    for (int r = 0; r < R; ++r) {  // one load per row of the wave
                        const int jr = wrap1(j0 + R * wr + r, ny);
                        if (lane < NXI * 7)
                            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                                rxi, (__attribute__((address_space(3))) void*)(lds + OFF_XIB + (par * TY + R * wr + r) * 64),
                                4, (int)(4u * lane), (int)(4u * (unsigned)(((long long)tk * ny + jr) * (NXI * 7))), 0, 0);
                    }
                }
            };
            // rows of directions 0, 1, 2 of target tk, staged by the group B wave of each row
            // for the group A hand-over (no z dependence)
            auto fetch012 = [&](int tk, int par) {
                if constexpr (G == 1 && W) {
                    // one 16-byte load: direction 0 at the nodes, then 1 and 2 with the x halo
                    // (FP32: granules 0-15, 16-33, 34-51; 16-bit: 0-7, 8-17, 18-27); the
                    // remaining lanes stay idle
                    constexpr int n012 = TX / LY::GV + 2 * (HXW / LY::GV);  // granules of the three rows
                    static_assert(4 * TX == P_ROW || R == 1, "lanes of row rl land at row rl's slot");
                    float* P = lds + OFF_P012 + (par * TY + R * wr) * P_ROW;
                    const unsigned lv = SZ * (unsigned)tk * (unsigned)plane;
                    if constexpr (BIG) {
                        // one load per direction, each row's granules where the single load puts them
                        constexpr int n0 = TX / LY::GV, n12 = HXW / LY::GV;
                        if (tx < n0) lds_dma16(src(0), P, p16q[0], lv);
                        if (tx < n12) lds_dma16(src(1), P + 4 * n0, p16q[1], lv);
                        if (tx < n12) lds_dma16(src(2), P + 4 * (n0 + n12), p16q[2], lv);
                    } else {
                        if (tx < n012) lds_dma16(rfo, P, p16, lv);
                    }
                } else if constexpr (G == 1) {
                    const unsigned lv = SZ * (unsigned)tk * (unsigned)plane;
                    float* P = lds + OFF_P012 + (par * TY + ty) * P_ROW;
                    const unsigned o = lv + yrow[3 + ty];
                    dma(src(0), xn, qoff(0) + o, P);
                    dma(src(1), xh, qoff(1) + o, P + TX);
                    dma(src(1), xh6, qoff(1) + o, P + TX + 6);
                    dma(src(2), xh, qoff(2) + o, P + TX + HX);
                    dma(src(2), xh6, qoff(2) + o, P + TX + HX + 6);
                }
            };
            // hand-over targets of the next two steps (k0 where a step hands nothing over: the
            // load keeps the per-step count fixed and fills a slot nobody reads)
            auto lookahead = [&](const Sched& from, int& h1, int& h2) {
                Sched n = from;
                bool a_;
                sched_step(n, k1, nz, a_, h1);
                sched_step(n, k1, nz, a_, h2);
                h1 = h1 >= 0 ? h1 : k0;
                h2 = h2 >= 0 ? h2 : k0;
            };
            {
                // the two steps before step 0: level L0, the xi weights and rows of 0-2 of
                // step 0, then those of step 1 and level L0 + 1
                Sched n = sc;
                bool a_;
                int h0, h1;
                sched_step(n, k1, nz, a_, h0);
                sched_step(n, k1, nz, a_, h1);
                fetch_xi(h0 >= 0 ? h0 : k0, 0);
                fetch012(h0 >= 0 ? h0 : k0, 0);
                fetch_xi(h1 >= 0 ? h1 : k0, 1);
                fetch012(h1 >= 0 ? h1 : k0, 1);
            }
            fetch(min(L0 + 1, L1), 1);
            int outstanding = ntab + nbatch;  // loads issued in the step before
            // step s has parity P = s % 2: the loop below alternates the two forms
            auto step = [&](auto pc) -> bool {
                constexpr int P = decltype(pc)::value;
                if (!sched_step(sc, k1, nz, adv, ho)) return false;
                const int x = sc.top;  // level appended in this step (if adv)
                // this wave's share of level x and of the rows of directions 0-2 and xi weights
                // staged for this step are in LDS
                wait_staged(outstanding);
                if (adv) decode(b);
                if (adv) decode16(b, db);
                lds_barrier();  // level x staged by all waves; level x-1 and hand-over s-2 no longer read
                // level x (steps that append): staged, or its decoded copy
                const float* S = DEC ? lds + LY::OFF_DEC + db * LY::DEC_SZ : lds + b * STAGE;
                {
                    int h1, h2;
                    lookahead(sc, h1, h2);
                    const int s1 = (s + 1) % NBUF, s2 = (s + 2) % NBUF;
                    fetch_xi(h2, s2);   // read by the ring waves in step s+2
                    fetch012(h2, s2);   // read by the group A waves in step s+2
                }
                if (adv) {
                    // three raw buffers: the buffer after the next holds level x-1, filtered in
                    // the last append step; two (decoded levels): level x's raw buffer, decoded
                    // by every wave before the barrier above
                    const int bn = b + 2 >= NRAW ? b + 2 - NRAW : b + 2;
                    fetch(min(x + 2, L1), bn);
                    b = b + 1 == NRAW ? 0 : b + 1;
                }
                outstanding = ntab + (adv ? nbatch : 0);
                float nv[8];  // this step's filtered values of the group's directions
                if (adv) {
                    // horizontal filters of this group's directions.
                    // Staged value at offset o of level x: FP32 as stored, or 16-bit codes
                    // (16-byte staging) converted to FP32 before their direction's scale (in
                    // the decoded copy, or here where there is none); the scale multiplies the
                    // filtered value (the filters are linear).
                    constexpr bool code16 = W && k16;
                    constexpr bool raw16 = code16 && !DEC;
                    const unsigned short* S16 = (const unsigned short*)S;
                    auto sv = [&](int o) -> float {
                        if constexpr (raw16)
                            return dec16<ST>(S16[o]);
                        else
                            return S[o];
                    };
                    auto scl = [&](float v, int q) -> float {
                        if constexpr (code16)
                            return v * isc_of(q);
                        else
                            return v;
                    };
                    auto xf = [&](int o, const float* wv) {
                        float sum = sv(o) * wv[0];
#pragma unroll
                        for (int l = 1; l < 7; ++l) sum = sum + sv(o + l) * wv[l];
                        return sum;
                    };
                    auto xfp = [&](const float* src, const float* wv) {
                        float sum = src[0] * wv[0];
#pragma unroll
                        for (int l = 1; l < 7; ++l) sum = sum + src[l] * wv[l];
                        return sum;
                    };
                    auto yf = [&](int o, const float* wv) {
                        float sum = sv(o) * wv[0];
#pragma unroll
                        for (int n = 1; n < 7; ++n) sum = sum + sv(o + n * TX) * wv[n];
                        return sum;
                    };
                    if constexpr (G == 0) {
                        // y pass of 7-10 on this wave's row into its own LDS area (read back
                        // by this wave only: no barrier needed), over the 70 columns the x pass
                        // reads (from x = i0 - 3): 64 per direction, then the last 6 of all
                        // four directions in one pass of 24 lanes
                        float* YP = lds + OFF_YP + ty * 4 * HXW;
#pragma unroll
                        for (int d = 0; d < 4; ++d) {
                            const float* wyy = wy[ysel(7 + d)];
                            const int c = XO + tx;
                            const int o = OFF_XY + d * XY_SZ + ty * HXW + c;
                            float sum = sv(o) * wyy[0];
#pragma unroll
                            for (int n = 1; n < 7; ++n) sum = sum + sv(o + n * HXW) * wyy[n];
                            YP[d * HXW + c] = sum;
                        }
                        static_assert(ysel(7) == ysel(8) && ysel(9) == ysel(10), "7, 8 and 9, 10 share y weights");
                        if (lane < 24 * R) {  // row R wr + lane / 24
                            const int e = lane % 24, row = R * wr + lane / 24;
                            const int d = e / 6, c = XO + TX + e % 6;
                            const bool first = d < 2;
                            const int o = OFF_XY + d * XY_SZ + row * HXW + c;
                            float sum = sv(o) * (first ? wy[ysel(7)][0] : wy[ysel(9)][0]);
#pragma unroll
                            for (int n = 1; n < 7; ++n)
                                sum = sum + sv(o + n * HXW) * (first ? wy[ysel(7)][n] : wy[ysel(9)][n]);
                            lds[OFF_YP + row * 4 * HXW + d * HXW + c] = sum;
                        }
                        if constexpr (SW) {
                            nv[0] = scl(sv(OFF_Z + (TY + ty) * TX + tx), 6);                           // 6 at the node
                            nv[1] = scl(xf(OFF_X + 2 * X_SZ + ty * HXW + XO + tx, wx[xsel(13)]), 13);  // 13: x filter
                        } else {
                            nv[0] = scl(sv(OFF_Z + ty * TX + tx), 5);         // 5 at the node (no filter)
                            nv[1] = scl(sv(OFF_Z + (TY + ty) * TX + tx), 6);  // 6
                        }
#pragma unroll
                        for (int d = 0; d < 4; ++d)  // 7-10: x pass over the y-passed row
                            nv[2 + d] = scl(xfp(YP + d * HXW + XO + tx, wx[xsel(7 + d)]), 7 + d);
#pragma unroll
                        for (int d = 0; d < 2; ++d)  // 11, 12: x filter
                            nv[6 + d] = scl(xf(OFF_X + d * X_SZ + ty * HXW + XO + tx, wx[xsel(11 + d)]), 11 + d);
                    } else {
#pragma unroll
                        for (int d = 0; d < 2; ++d)  // 3, 4: y filter
                            nv[d] = scl(yf(OFF_Y + d * Y_SZ + ty * TX + tx, wy[ysel(3 + d)]), 3 + d);
                        if constexpr (SW) {
                            nv[2] = scl(sv(OFF_Z + ty * TX + tx), 5);                                  // 5 at the node
                            nv[3] = scl(xf(OFF_X + 3 * X_SZ + ty * HXW + XO + tx, wx[xsel(14)]), 14);  // 14: x filter
                        } else {
#pragma unroll
                            for (int d = 2; d < 4; ++d)  // 13, 14: x filter
                                nv[d] = scl(xf(OFF_X + d * X_SZ + ty * HXW + XO + tx, wx[xsel(11 + d)]), 11 + d);
                        }
#pragma unroll
                        for (int d = 2; d < 6; ++d)  // 15-18: y filter
                            nv[2 + d] = scl(yf(OFF_Y + d * Y_SZ + ty * TX + tx, wy[ysel(qy(d))]), qy(d));
                    }
                }
                if (adv) db ^= 1;
                // hand-over of target ho from the window of the last append (parity S)
                auto handover = [&](auto wc) {
                    constexpr int S = decltype(wc)::value;
                    // xi weights of the target (wave-uniform LDS reads, staged in step s-1)
                    const float* XI = lds + OFF_XIB + ((s % NBUF) * TY + ty) * 64;
                    float* HO = lds + OFF_HO + (s & 1) * HO_SZ;
                    auto win = [&](int d, int m) -> float { return h[d][m + (S == 0 ? 1 : 0)]; };  // S = P
#pragma unroll
                    for (int d = 0; d < 8; ++d) {
                        const int q = gq(G, d, SW);
                        const float* wz = XI + xifam(q) * 7;
                        float sum = win(d, 0) * wz[0];
#pragma unroll
                        for (int m = 1; m < 7; ++m) sum = sum + win(d, m) * wz[m];
                        HO[q * (TX * TY) + col] = sum;
                    }
                    if constexpr (G == G012) {
                        // directions 0, 1, 2 of the target from its rows, staged two steps ahead by
                        // group B (its wait for them precedes the barrier of this step)
                        float f0, f1, f2;
                        reduce012(lds + OFF_P012 + ((s % NBUF) * TY + ty) * P_ROW, wx[0], wx[1], f0, f1, f2);
                        HO[col] = f0;
                        HO[TX * TY + col] = f1;
                        HO[2 * TX * TY + col] = f2;
                    }
                };
                // after step s the window is that of parity s % 2: an appending step writes
                // its slot (odd: after moving the pairs down), a step that appends nothing
                // moves the window by one slot instead (the first and last levels of the
                // column only)
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    if (adv) {
                        if constexpr (P == 1) {
#pragma unroll
                            for (int m = 0; m < 6; ++m) h[d][m] = h[d][m + 2];
                            h[d][6] = nv[d];
                        } else {
                            h[d][7] = nv[d];
                        }
                    } else {
                        if constexpr (P == 1) {
#pragma unroll
                            for (int m = 0; m < 7; ++m) h[d][m] = h[d][m + 1];
                        } else {
#pragma unroll
                            for (int m = 7; m > 0; --m) h[d][m] = h[d][m - 1];
                        }
                    }
                }
                if (ho >= 0) handover(std::integral_constant<int, P>{});
                ++s;
                return true;
            };
            // the ring starts empty (zeros), as after an odd step
            for (;;) {
                if (!step(std::integral_constant<int, 0>{})) break;
                if (!step(std::integral_constant<int, 1>{})) break;
            }
            // the final step, in which the collision waves take the last target: no staging
            // load may still write LDS when the workgroup retires
            __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
            lds_barrier();
        };
        if (w < NWR)
            ring(std::integral_constant<int, 0>{});
        else
            ring(std::integral_constant<int, 1>{});
    } else {
        // ------------------------------------------------------------- collision waves
        // In step s they collide the target handed over in step s-1, while the ring waves
        // filter and hand over the next one.
        const int cw = wave - NRW;
        const int ty = R * cw + rl;  // row of this lane
        const int col = ty * TX + tx;
        const int j = j0 + ty;
        const int jw = j < ny ? j : j % ny;
        const unsigned own = (unsigned)(jw * nx + iw);
        const unsigned own_b = SZ * own;
        const unsigned own4 = 4u * own;
        const unsigned row = (unsigned)(jw * nx);
        // Scalars and tables that no kernel writes during the step are read through the
        // constant address space (scalar loads): a vector load here would be waited for
        // together with the next target's loads issued just before it.
        using cdouble = const __attribute__((address_space(4))) double;
        const float rho_mod = (float)*(cdouble*)a.rho_modify;
        const float force = (float)*(cdouble*)a.force;
        const unsigned fbytes = 4u * (unsigned)N;
        const auto rfn = rsrc_of(a.fn, BIG ? nbq : NQ * nbq);
        auto dst = [&](int q) {
            if constexpr (BIG)
                return wide_rsrc(a.fn, (unsigned long long)q * nbq64, nbq);
            else
                return rfn;
        };
        // rho, u, v, w are consecutive arrays of one allocation (checked at launch)
        const auto rmac = rsrc_of(a.rho, BIG ? fbytes : 4 * fbytes);
        // rho, u, v, w: one resource, or with wide addressing one per field
        auto macdst = [&](int f) {
            if constexpr (BIG)
                return wide_rsrc(a.rho, (unsigned long long)f * fbytes, fbytes);
            else
                return rmac;
        };
        auto macoff = [&](int f) { return BIG ? 0u : (unsigned)f * fbytes; };
        // Bulk-flux probe (AccumulateUbulk, evolution.h:826-840) fused into the step.  The
        // reference samples v of the previous step on the probe plane before each step; here
        // the waves of the probe row sum v dx_i dz_k of the values they store, per lane over
        // the chunk's levels and then over the wave, into this step's slot (chunk, x tile),
        // and the next step adds the slots of this one to the accumulator in a fixed order.
        // Lanes outside the grid (partial tiles) carry weight 0.  The global loads below
        // precede every store of the wave.
        const bool probe = a.flux_part != nullptr && j == a.flux_plane;
        // a wave holding the probe row (its other rows' lanes add zeros)
        const bool probe_wave = a.flux_part != nullptr && a.flux_plane >= j0 + R * cw && a.flux_plane < j0 + R * cw + R;
        double pdx = 0.0, psum = 0.0;
        if (probe) pdx = i < nx ? a.flux_dx[iw] : 0.0;
        if (a.flux_part != nullptr && wk == 0 && cw == 0 && lane == 0) {
            // the previous step's partials (written by the previous launch), in slot order
            const double* prev = a.flux_part + (long long)(1 - a.flux_par) * a.flux_nslot;
            double t = 0.0;
            for (int e = 0; e < a.flux_nslot; ++e) t += prev[e];
            a.flux_acc[0] += t;
        }
        // 16-bit storage: the step count that dithers the stochastic rounding of the stores
        // (slot flux_par, written by the previous launch); one lane of the launch writes the next
        unsigned sr_step = 0;
        if constexpr (k16) {
            const double ctr = ((cdouble*)a.step_ctr)[a.flux_par];
            sr_step = (unsigned)(unsigned long long)ctr;
            if (wk == 0 && cw == 0 && lane == 0) a.step_ctr[1 - a.flux_par] = ctr + 1.0;
        }
        // walls, collision and stores of target level tk.  Lanes outside the grid (partial
        // tiles) hold the wrapped column or row, whose values they compute bit for bit like
        // its owner, so their stores duplicate the owner's: the stores need no lane mask,
        // and every step issues the same vector-memory sequence.
        auto collide_store = [&](float (&F)[NQ], int tk) {
            const long long node = (long long)tk * plane + own;
            if (tk <= 1 || tk == nz - 1) {
                int xo[7], yo[7];
#pragma unroll
                for (int l = 0; l < 7; ++l) {
                    xo[l] = wrap_index(iw - 3 + l, nx);
                    yo[l] = wrap_index(jw - 3 + l, ny) * nx;
                }
                // 16-bit storage: wall reads through the buffer resource (no 64-bit address per array
                // held in scalar registers: BF16 152.9 -> 148.9 us, and with the literal scales FP16
                // 140.8 -> 139.9 us at 128x256x128); FP32 times the same either way and keeps the
                // global-pointer reads
                if constexpr (k16) {
                    auto pop = [&](int q) {
                        return [&, q](long long n) { return IO::ld(src(q), SZ * (unsigned)n, qoff(q), isc_of(q)); };
                    };
                    apply_walls_with<ST>(F, a, pop, node, iw, jw, tk, xo, yo, plane);
                } else {
                    apply_walls<ST>(F, a, node, iw, jw, tk, xo, yo, plane);
                }
            }
            // LES filter width of the row (scalar loads: one per row of the wave)
            const auto les_tab = (const __attribute__((address_space(4))) float*)a.t.les_delta;
            float delta = 0.f;
            if (a.les) {
                if constexpr (R == 1) {
                    delta = les_tab[(long long)tk * ny + jw];
                } else {
                    const float d0 = les_tab[(long long)tk * ny + wrap1(j0 + R * cw, ny)];
                    const float d1 = les_tab[(long long)tk * ny + wrap1(j0 + R * cw + 1, ny)];
                    delta = rl ? d1 : d0;
                }
            }
            float rho, ux, uy, uz, tau_eff;
            collide_pairs<true>(F, rho_mod, (float)a.dt, force, delta, a.les, (float)a.tau, (float)a.cs, rho, ux, uy, uz,
                                tau_eff, (float)(2.0 * a.tau - 1.0));
            if (probe) psum += ((double)uy * pdx) * ((const __attribute__((address_space(4))) double*)a.flux_dz)[tk];
            const unsigned lv = SZ * (unsigned)tk * (unsigned)plane;
            const unsigned lf = 4u * (unsigned)tk * (unsigned)plane;
            if constexpr (k16) {
                const unsigned h = sr_hash((unsigned)node, sr_step);
#pragma unroll
                for (int q = 0; q < NQ; ++q) IO::st_sr(dst(q), own_b, qoff(q) + lv, F[q], scl_of(q), h + sr_lane(q));
            } else {
#pragma unroll
                for (int q = 0; q < NQ; ++q) IO::st(dst(q), own_b, qoff(q) + lv, F[q], scl_of(q));
            }
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(rho), macdst(0), (int)own4, (int)(lf + macoff(0)), kStoreAux);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(ux), macdst(1), (int)own4, (int)(lf + macoff(1)), kStoreAux);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(uy), macdst(2), (int)own4, (int)(lf + macoff(2)), kStoreAux);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(uz), macdst(3), (int)own4, (int)(lf + macoff(3)), kStoreAux);
        };
        Sched sc{L0 - 1, k0};
        bool adv;
        int ho;
        int s = 0;
        // warm-up: the steps up to the first hand-over
        for (;;) {
            sched_step(sc, k1, nz, adv, ho);
            lds_barrier();
            ++s;
            if (ho >= 0) break;
        }
        int pend = ho;  // target handed over in step s-1
        // One target per step.  The collision waves issue no global loads: a wait for a load
        // would also wait for every store issued before it (one in-order counter), so their
        // inputs all come through LDS.
        for (;;) {
            const bool more = sched_step(sc, k1, nz, adv, ho);
            lds_barrier();  // hand-over of pend (step s-1) in LDS
            float F[NQ];
            const float* HO = lds + OFF_HO + ((s - 1) & 1) * HO_SZ;
#pragma unroll
            for (int q = 0; q < NQ; ++q) F[q] = HO[q * (TX * TY) + col];
            collide_store(F, pend);
            if (!more) break;  // final step: the ring has appended and handed over everything
            pend = ho;
            ++s;
        }
        if (probe_wave) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) psum += __shfl_xor(psum, off);
            if (lane == 0) a.flux_part[(long long)a.flux_par * a.flux_nslot + chunk * tg.ntx + tile % tg.ntx] = psum;
        }
    }
}

template <Store ST, bool W, int TXW, bool BIG>
__global__ __launch_bounds__(Lay<W>::NTH) __attribute__((amdgpu_waves_per_eu(3))) void k_collide_tiled(StepArgs<ST> a,
                                                                                                          TiledGeom tg)
{
    tiled_body<ST, W, TXW, BIG>(a, tg, (int)blockIdx.x);
}

// Several solves advanced by one launch (one step of each): solve e owns blocks
// [block0[e], block0[e + 1]), each range starting at a multiple of 8 so that a block runs on
// the XCD its solve's work mapping expects; a solve whose step limit within the current
// chunk is reached takes no part.  The per-solve arguments stay in the kernel-argument
// segment, read with scalar loads like those of the single-solve kernel.
constexpr int kGroupMax = 8;
template <Store ST>
struct GroupArgs {
    StepArgs<ST> a[kGroupMax];
    TiledGeom tg[kGroupMax];
    int block0[kGroupMax + 1];
    int limit[kGroupMax];  // steps of the chunk the solve takes
    int n;                 // solves
    int step;              // index of this launch within the chunk
};
static_assert(sizeof(GroupArgs<Store::F32>) <= 4096, "grouped kernel arguments exceed 4 KiB");

static_assert(Lay<true, 4, 32>::NTH == Lay<true>::NTH && Lay<true, 2, 32>::NTH == Lay<true>::NTH,
              "both tile widths run 12 waves");
template <Store ST, bool W, int TXW>
__global__ __launch_bounds__(Lay<W>::NTH) __attribute__((amdgpu_waves_per_eu(3))) void k_collide_tiled_group(
    GroupArgs<ST> g)
{
    const int b = (int)blockIdx.x;
    int e = 0;
    while (e + 1 < g.n && b >= g.block0[e + 1]) ++e;
    if (g.step >= g.limit[e] || b - g.block0[e] >= g.tg[e].nblocks) return;
    tiled_body<ST, W, TXW>(g.a[e], g.tg[e], b - g.block0[e]);
}

bool tiled_wide_staging(const LaunchGeom& g, Store store);

// Tile width of a solve with 16-byte staging: 32 columns where that leaves fewer lanes idle
// than 64-column tiles (nx = 16, 32, 80, 96, ..., and for FP32 also 20, 36, ...), else 64.
// Fixed by the grid, so a solve always runs the same compiled kernel.
int tiled_tile_width(const LaunchGeom& g, Store store)
{
    if (!tiled_wide_staging(g, store)) return 64;
    const int lanes64 = (g.nx + 63) / 64 * 64, lanes32 = (g.nx + 31) / 32 * 32;
    return lanes32 < lanes64 ? 32 : 64;
}

TiledGeom tiled_geom_of(const LaunchGeom& g, Store store)
{
    if (!tiled_wide_staging(g, store)) return tiled_geom(g.nx, g.ny, g.nz, 64, Lay<false>::TY, Lay<false>::RESIDENT);
    if (tiled_tile_width(g, store) == 32)
        return tiled_geom(g.nx, g.ny, g.nz, 32, Lay<true, 4, 32>::TY, Lay<true, 4, 32>::RESIDENT);
    return tiled_geom(g.nx, g.ny, g.nz, 64, Lay<true>::TY, Lay<true>::RESIDENT);
}

bool tiled_supported(const LaunchGeom& g, Store store, bool fp64)
{
    // buffer resources address a population array, a field or the xi table with 32-bit byte
    // offsets (with wide addressing: one array per resource)
    const long long sz = store == Store::F32 ? 4 : 2;
    return !fp64 && store != Store::F64 && g.nx >= 8 && g.ny >= 8 && g.nz >= 8 && g.qs * sz < (1LL << 32) &&
           g.n * 4 < (1LL << 32) && (long long)g.nz * g.ny * NXI * 7 * 4 < (1LL << 32);
}

bool tiled_needs_wide(const LaunchGeom& g, Store store)
{
    // one resource for a buffer's 19 arrays, and one for rho, u, v, w
    const long long sz = store == Store::F32 ? 4 : 2;
    return NQ * g.qs * sz >= (1LL << 32) || 4 * g.n * 4 >= (1LL << 32);
}

void tiled_tile_shape(const LaunchGeom& g, Store store, int* cols, int* rows)
{
    if (!tiled_wide_staging(g, store)) {
        *cols = 64;
        *rows = Lay<false>::TY;
        return;
    }
    *cols = tiled_tile_width(g, store);
    *rows = *cols == 32 ? Lay<true, 4, 32>::TY : Lay<true>::TY;
}

int tiled_flux_slots(const LaunchGeom& g, Store store)
{
    const TiledGeom t = tiled_geom_of(g, store);
    return t.nchunks * t.ntx;
}

template <Store ST, bool W, int TXW = 64, bool BIG = false>
static hipError_t launch_t(hipStream_t st, const StepLaunch& L)
{
    StepArgs<ST> a = make_step_args<ST>(L);
    using LY = Lay<W, W ? (ST == Store::F32 ? 4 : 2) : 4, TXW>;
    TiledGeom tg = tiled_geom(L.g.nx, L.g.ny, L.g.nz, TXW, LY::TY, LY::RESIDENT);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_collide_tiled<ST, W, TXW, BIG>), dim3(tg.nblocks), dim3(LY::NTH), 0, st, a,
                       tg);
    return hipGetLastError();
}

// the tiled kernel with one resource per population array and per field (wide addressing)
template <Store ST>
static hipError_t launch_wide(hipStream_t st, const StepLaunch& L)
{
    if (!tiled_wide_staging(L.g, L.store)) return launch_t<ST, false, 64, true>(st, L);
    return tiled_tile_width(L.g, L.store) == 32 ? launch_t<ST, true, 32, true>(st, L) : launch_t<ST, true, 64, true>(st, L);
}

bool tiled_wide_staging(const LaunchGeom& g, Store store)
{
    // whole 16-byte granules of a row (4 FP32 or 8 16-bit values) never straddle the
    // periodic wrap in x
    return store == Store::F32 ? g.nx % 4 == 0 : g.nx % 8 == 0;
}

bool tiled_groupable(const LaunchGeom& g, Store store)
{
    return store == Store::F32 && tiled_supported(g, store, false) && tiled_wide_staging(g, store) &&
           !tiled_needs_wide(g, store);
}

// The members of one tile width go into one launch: a group of mixed widths takes two
// launches per step, on st one after the other, or with `fork` the 32-column one on a second
// stream so that the two run together (the next step waits for both).
hipError_t launch_tiled_group(hipStream_t st, const StepLaunch* L, const int* limit, int n, int step,
                              const GroupFork* fork)
{
    if (n <= 0 || n > kGroupMax) return hipErrorInvalidValue;
    for (int e = 0; e < n; ++e) {
        if (L[e].store != Store::F32 || !tiled_wide_staging(L[e].g, L[e].store)) return hipErrorInvalidValue;
        const size_t fb = (size_t)L[e].g.n * 4;
        const char* r = (const char*)L[e].rho;
        if ((const char*)L[e].u != r + fb || (const char*)L[e].v != r + 2 * fb || (const char*)L[e].w != r + 3 * fb)
            return hipErrorInvalidValue;
    }
    bool mixed = false;
    for (int e = 1; e < n; ++e) mixed = mixed || tiled_tile_width(L[e].g, L[e].store) != tiled_tile_width(L[0].g, L[0].store);
    const bool forked = mixed && fork != nullptr;
    if (forked) {
        hipError_t e = hipEventRecord(fork->fork, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(fork->st2, fork->fork, 0);
        if (e != hipSuccess) return e;
    }
    for (const int width : {64, 32}) {
        const hipStream_t sw = forked && width == 32 ? fork->st2 : st;
        GroupArgs<Store::F32> g{};
        int blocks = 0, m = 0;
        for (int e = 0; e < n; ++e) {
            if (tiled_tile_width(L[e].g, L[e].store) != width) continue;
            g.a[m] = make_step_args<Store::F32>(L[e]);
            g.tg[m] = tiled_geom_of(L[e].g, L[e].store);
            g.block0[m] = blocks;
            g.limit[m] = limit[e];
            blocks += g.tg[m].nblocks;  // a multiple of 8
            ++m;
        }
        if (m == 0) continue;
        g.block0[m] = blocks;
        g.n = m;
        g.step = step;
        if (width == 64)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_collide_tiled_group<Store::F32, true, 64>), dim3(blocks),
                               dim3(Lay<true>::NTH), 0, sw, g);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_collide_tiled_group<Store::F32, true, 32>), dim3(blocks),
                               dim3(Lay<true, 4, 32>::NTH), 0, sw, g);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (forked) {
        hipError_t e = hipEventRecord(fork->join, fork->st2);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, fork->join, 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_tiled(hipStream_t st, const StepLaunch& L)
{
    // the kernel addresses rho, u, v, w through one buffer resource
    const char* r = (const char*)L.rho;
    const size_t fb = (size_t)L.g.n * 4;
    if ((const char*)L.u != r + fb || (const char*)L.v != r + 2 * fb || (const char*)L.w != r + 3 * fb)
        return hipErrorInvalidValue;
    if (L.wide_addr || tiled_needs_wide(L.g, L.store)) {
        switch (L.store) {
        case Store::F32: return launch_wide<Store::F32>(st, L);
        case Store::BF16: return launch_wide<Store::BF16>(st, L);
        case Store::FP16: return launch_wide<Store::FP16>(st, L);
        default: return hipErrorNotSupported;
        }
    }
    switch (L.store) {
    case Store::F32: {
        // a solve that a group can take runs the grouped kernel alone, so that its steps are
        // the same compiled code whether it runs alone or in a group
        if (tiled_groupable(L.g, L.store)) {
            const int one = 1;
            return launch_tiled_group(st, &L, &one, 1, 0);
        }
        return launch_t<Store::F32, false>(st, L);
    }
    case Store::BF16:
        if (!tiled_wide_staging(L.g, L.store)) return launch_t<Store::BF16, false>(st, L);
        return tiled_tile_width(L.g, L.store) == 32 ? launch_t<Store::BF16, true, 32>(st, L)
                                                    : launch_t<Store::BF16, true>(st, L);
    case Store::FP16:
        if (!tiled_wide_staging(L.g, L.store)) return launch_t<Store::FP16, false>(st, L);
        return tiled_tile_width(L.g, L.store) == 32 ? launch_t<Store::FP16, true, 32>(st, L)
                                                    : launch_t<Store::FP16, true>(st, L);
    default: return hipErrorNotSupported;
    }
}

}  // namespace islbm
