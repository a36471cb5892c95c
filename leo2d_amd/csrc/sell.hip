// sell.hip -- Chebyshev step on SELL-4-sigma row groups with the neighbour rows
// of each row tile staged in LDS.
//
// One launch = one step T_{j+1} = alpha H T_j - beta T_{j-1} for a block of R
// vectors (row-major [n][R]) fused with the per-vector partial sums of the
// moment reductions, like the CSR step kernels in kpm.hip.  It replaces the
// same loop of the reference (locality.cpp:6002-6041 / 6089-6127 calling
// SpMatrix::vectorMultiply, spmatrix.cpp:346-360 / 395-410).
//
// Why: a CSR step gathers one whole row of T_j (R x 16 B = 1 KiB at R = 64
// complex FP64) per stored entry, about 32 times per output row on bilayer
// graphene, all through the vector-memory path (L2 -> L1 -> lanes).  Here a
// workgroup owns a tile of up to 64 spatially consecutive rows; for each chunk
// of C vectors (C x element = 256 B) it copies the tile's distinct neighbour
// row chunks (the tile's "union", about 4 per output row) once into LDS with
// direct global->LDS loads, and every entry then reads its operand from LDS.
//
// Layout of the matrix (built on the host once per sparsity pattern):
//  * tiles: consecutive rows [tile_row0[t], tile_row0[t+1]) of the spatially
//    ordered matrix; the tile's union is ucols[tile_u0[t] ...], own rows
//    first (union index u = row - tile_row0[t]) so T_j of an own row is in LDS;
//  * SELL-4-sigma groups: the tile's rows sorted by length (sigma = the tile)
//    and cut into groups of 4; group g holds the rows grow[4g .. 4g+3] (-1 for
//    an empty slot) and glen[g] steps (the longest row, rounded up to 4);
//    entry e of slot s is at goff[g] + 4 e + s (padding: value 0, index 0);
//  * each entry is (eu = local union index (uint16), ev = value).
//
// Lanes: a wavefront runs one group at a time; 16-lane DPP row s is matrix row
// grow[4g+s] and lane k of it is vector k (or 2k, 2k+1 for FP32 vectors) of the
// chunk.  Lane (s, k) keeps entries k, k+16, k+32, k+48 of its row in
// registers for all chunks of the tile; step e broadcasts entry e inside each
// DPP row (row_newbcast), so the four rows advance together and every row's
// products are added in CSR order.  One LDS read of 16 B per lane per step:
// the four rows' operand chunks are four 256-byte runs.
//
// Reductions are deterministic (fixed tile order per workgroup, fixed lane
// and wavefront folds, no floating-point atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernels.h"
#include "kpm_common.h"

namespace leo2d {
namespace {


typedef float f4v __attribute__((ext_vector_type(4)));

// one 16-byte LDS read of the VPL vector elements of a lane
template <typename XT, int VPL>
__device__ inline void lds_ld(XT (&x)[VPL], const XT* p)
{
    if constexpr (VPL == 1) {
        x[0] = *p;
    } else if constexpr (std::is_same<XT, double>::value) {
        typedef double d2v __attribute__((ext_vector_type(2)));
        const d2v q = *reinterpret_cast<const d2v*>(p);
        x[0] = q.x;
        x[1] = q.y;
    } else {
        const f4v q = *reinterpret_cast<const f4v*>(p);
        x[0] = make_float2(q.x, q.y);
        x[1] = make_float2(q.z, q.w);
    }
}

// T_{j-1} rows in and T_{j+1} rows out of the copy-wavefront step are streamed:
// each lane moves its VPL elements (16 or 8 bytes) once per step, so they are
// read and written non-temporally and leave L2 to the staged rows of T_j,
// which neighbouring tiles read again (c1: 1-1.5 % per step, same results)
template <typename XT, int VPL>
__device__ inline void ld_stream(XT (&x)[VPL], const XT* p)
{
    static_assert(VPL * sizeof(XT) == 16 || VPL * sizeof(XT) == 8, "16 or 8 bytes per lane");
    if constexpr (VPL * sizeof(XT) == 16) {
        typedef int i4v __attribute__((ext_vector_type(4)));
        const i4v q = __builtin_nontemporal_load(reinterpret_cast<const i4v*>(p));
        __builtin_memcpy(&x[0], &q, 16);
    } else {
        typedef int i2v __attribute__((ext_vector_type(2)));
        const i2v q = __builtin_nontemporal_load(reinterpret_cast<const i2v*>(p));
        __builtin_memcpy(&x[0], &q, 8);
    }
}
template <typename XT, int VPL>
__device__ inline void st_stream(XT* p, const XT (&x)[VPL])
{
    static_assert(VPL * sizeof(XT) == 16 || VPL * sizeof(XT) == 8, "16 or 8 bytes per lane");
    if constexpr (VPL * sizeof(XT) == 16) {
        typedef int i4v __attribute__((ext_vector_type(4)));
        i4v q;
        __builtin_memcpy(&q, &x[0], 16);
        __builtin_nontemporal_store(q, reinterpret_cast<i4v*>(p));
    } else {
        typedef int i2v __attribute__((ext_vector_type(2)));
        i2v q;
        __builtin_memcpy(&q, &x[0], 8);
        __builtin_nontemporal_store(q, reinterpret_cast<i2v*>(p));
    }
}

// entries k, k+16, ..., k+16(QS-1) of the lane's row (union index, value)
template <typename HT, int QS>
struct GroupRegs {
    int eu[QS];
    HT ev[QS];
};

// operands of steps 4B .. 4B+3 (entry e is held by lane e % 16 of register e / 16)
template <int B, typename HT, typename XT, int VPL, int QS>
__device__ inline void ld_block(XT (&x)[4][VPL], const GroupRegs<HT, QS>& g, const XT* xs, int koff)
{
    {
        const int o = bc16<(4 * B + 0) & 15>(g.eu[(4 * B + 0) >> 4]);
        lds_ld<XT, VPL>(x[0], xs + o * (16 * VPL) + koff);
    }
    {
        const int o = bc16<(4 * B + 1) & 15>(g.eu[(4 * B + 1) >> 4]);
        lds_ld<XT, VPL>(x[1], xs + o * (16 * VPL) + koff);
    }
    {
        const int o = bc16<(4 * B + 2) & 15>(g.eu[(4 * B + 2) >> 4]);
        lds_ld<XT, VPL>(x[2], xs + o * (16 * VPL) + koff);
    }
    {
        const int o = bc16<(4 * B + 3) & 15>(g.eu[(4 * B + 3) >> 4]);
        lds_ld<XT, VPL>(x[3], xs + o * (16 * VPL) + koff);
    }
}

// Complex products with the entry broadcast folded into the multiply-adds: the
// DPP form of v_fmac reads src0 from lane K of each 16-lane row, so one step is
// four (FP64, one vector) or eight (FP32, two vectors) fused multiply-adds, each
// accumulator in the order of mac(): re += hr xr, re -= hi xi, im += hr xi,
// im += hi xr.  The two-vector FP32 form runs the four accumulators' first
// products before their second ones, so no product waits on the one just
// issued (0.5-1.5 % per c1 FP32 step; the same order made the FP64 step 1.5 %
// slower, which keeps each accumulator's two products together).
//
// A DPP source needs two wait states after a VALU write of it.  The entry values
// are written only when a tile's groups are taken over, so one s_nop 1 before a
// group's products, tied to the value registers (they cannot move past it),
// covers every step of the group.  (An address formed by one DPP add in place
// of the broadcast and the shift-or measured slower: 872 against 823 us FP64.)
__device__ inline void pin_values(double& a) { asm volatile("s_nop 1" : "+v"(a)); }
__device__ inline void pin_values(float& a) { asm volatile("s_nop 1" : "+v"(a)); }
__device__ inline void pin_values(double2& a) { asm volatile("s_nop 1" : "+v"(a.x), "+v"(a.y)); }
__device__ inline void pin_values(float2& a) { asm volatile("s_nop 1" : "+v"(a.x), "+v"(a.y)); }
template <int K> struct DppMac;
#define LEO2D_DPP_MAC(K)                                                                                          \
    template <> struct DppMac<K> {                                                                                \
        __device__ static inline void f64(double& ax, double& ay, double hr, double hi, double xx, double xy)     \
        {                                                                                                          \
            asm("v_fmac_f64_dpp %0, %2, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f64_dpp %0, -%3, %5 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                \
                "v_fmac_f64_dpp %1, %2, %5 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f64_dpp %1, %3, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf"                       \
                : "+v"(ax), "+v"(ay)                                                                               \
                : "v"(hr), "v"(hi), "v"(xx), "v"(xy));                                                             \
        }                                                                                                          \
        __device__ static inline void f64r(double& a0, double& a1, double h, double x0, double x1)                \
        {                                                                                                          \
            asm("v_fmac_f64_dpp %0, %2, %3 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f64_dpp %1, %2, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf"                       \
                : "+v"(a0), "+v"(a1)                                                                               \
                : "v"(h), "v"(x0), "v"(x1));                                                                       \
        }                                                                                                          \
        __device__ static inline void f32x1(float& ax, float& ay, float hr, float hi, float xx, float xy)         \
        {                                                                                                          \
            asm("v_fmac_f32_dpp %0, %2, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f32_dpp %0, -%3, %5 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                \
                "v_fmac_f32_dpp %1, %2, %5 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f32_dpp %1, %3, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf"                       \
                : "+v"(ax), "+v"(ay)                                                                               \
                : "v"(hr), "v"(hi), "v"(xx), "v"(xy));                                                             \
        }                                                                                                          \
        __device__ static inline void f32(float& a0x, float& a0y, float& a1x, float& a1y, float hr, float hi,    \
                                          float x0x, float x0y, float x1x, float x1y)                              \
        {                                                                                                          \
            asm("v_fmac_f32_dpp %0, %4, %6 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f32_dpp %1, %4, %7 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f32_dpp %2, %4, %8 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f32_dpp %3, %4, %9 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f32_dpp %0, -%5, %7 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                \
                "v_fmac_f32_dpp %1, %5, %6 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                 \
                "v_fmac_f32_dpp %2, -%5, %9 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n\t"                \
                "v_fmac_f32_dpp %3, %5, %8 row_newbcast:" #K " row_mask:0xf bank_mask:0xf"                       \
                : "+v"(a0x), "+v"(a0y), "+v"(a1x), "+v"(a1y)                                                       \
                : "v"(hr), "v"(hi), "v"(x0x), "v"(x0y), "v"(x1x), "v"(x1y));                                       \
        }                                                                                                          \
    };
LEO2D_DPP_MAC(0) LEO2D_DPP_MAC(1) LEO2D_DPP_MAC(2) LEO2D_DPP_MAC(3)
LEO2D_DPP_MAC(4) LEO2D_DPP_MAC(5) LEO2D_DPP_MAC(6) LEO2D_DPP_MAC(7)
LEO2D_DPP_MAC(8) LEO2D_DPP_MAC(9) LEO2D_DPP_MAC(10) LEO2D_DPP_MAC(11)
LEO2D_DPP_MAC(12) LEO2D_DPP_MAC(13) LEO2D_DPP_MAC(14) LEO2D_DPP_MAC(15)
#undef LEO2D_DPP_MAC

template <int E, typename HT, typename XT, typename AT, int VPL, int QS>
__device__ inline void mac_step(AT (&acc)[VPL], const XT (&x)[VPL], const GroupRegs<HT, QS>& g)
{
    if constexpr (std::is_same<HT, double2>::value && VPL == 1) {
        const double2 h = g.ev[E >> 4];
        DppMac<E & 15>::f64(acc[0].x, acc[0].y, h.x, h.y, x[0].x, x[0].y);
    } else if constexpr (std::is_same<HT, float2>::value && VPL == 2) {
        const float2 h = g.ev[E >> 4];
        DppMac<E & 15>::f32(acc[0].x, acc[0].y, acc[1].x, acc[1].y, h.x, h.y, x[0].x, x[0].y, x[1].x, x[1].y);
    } else if constexpr (std::is_same<HT, float2>::value && std::is_same<XT, float2>::value && VPL == 1) {
        const float2 h = g.ev[E >> 4];
        DppMac<E & 15>::f32x1(acc[0].x, acc[0].y, h.x, h.y, x[0].x, x[0].y);
    } else if constexpr (std::is_same<HT, double>::value && std::is_same<XT, double>::value && VPL == 2) {
        DppMac<E & 15>::f64r(acc[0], acc[1], g.ev[E >> 4], x[0], x[1]);
    } else {
        const HT h = bc16<E & 15>(g.ev[E >> 4]);
#pragma unroll
        for (int v = 0; v < VPL; ++v) mac(acc[v], h, x[v]);
    }
}

// steps 4B .. L-1 of the group (L a multiple of 4, at most 16 QS); xa holds the
// operands of block B, the next block's are read before this block's products
// (one block past L is read and not used)
// PLO >= 0: the wavefront's issue priority drops to PLO when block 4 starts
template <int B, typename HT, typename XT, typename AT, int VPL, int QS, int PLO = -1>
__device__ inline void run_blocks(AT (&acc)[VPL], const GroupRegs<HT, QS>& g, const XT* xs, int koff, int L,
                                  const XT (&xa)[4][VPL])
{
    if constexpr (B < 4 * QS) {
        if (4 * B < L) {
            if constexpr (B == 4 && PLO >= 0) __builtin_amdgcn_s_setprio(PLO);
            XT xb[4][VPL];
            // unconditional (entries past L are index 0, a valid row): a branch here would
            // merge two LDS counts and make the wait below cover the next block too
            if constexpr (B + 1 < 4 * QS) ld_block<B + 1, HT, XT, VPL, QS>(xb, g, xs, koff);
            // keep those reads ahead of this block's products (the scheduler would sink them)
            __builtin_amdgcn_sched_barrier(0);
            mac_step<4 * B + 0, HT, XT, AT, VPL, QS>(acc, xa[0], g);
            mac_step<4 * B + 1, HT, XT, AT, VPL, QS>(acc, xa[1], g);
            mac_step<4 * B + 2, HT, XT, AT, VPL, QS>(acc, xa[2], g);
            mac_step<4 * B + 3, HT, XT, AT, VPL, QS>(acc, xa[3], g);
            run_blocks<B + 1, HT, XT, AT, VPL, QS, PLO>(acc, g, xs, koff, L, xb);
        }
    }
}

// s_waitcnt vmcnt(N) through the builtin, so the compiler's own wait insertion
// knows what has completed (expcnt and lgkmcnt left at their maxima, 7 and 15)
template <int N>
__device__ inline void wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt is six bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// per-wavefront-uniform metadata read through the constant address space: scalar
// loads, counted by lgkmcnt, so waiting for them never waits for the staging
// copies or the stores in flight (the arrays are not written during a step)
template <typename T>
__device__ inline T ldu(const T* p, int64_t i)
{
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Pipeline over (tile, chunk) pairs p of one workgroup, two LDS buffers:
//   top of pair p: wait for everything issued during pair p-1, barrier;
//     then issue T_{j-1} (and v) of this pair's rows, then the staging copies of
//     pair p+1 into the other buffer (always SPW per wavefront, a valid dummy
//     row when there is no pair p+1, so the wait counts stay static);
//   compute every group of the pair from LDS, outputs, then the stores;
//   then the loads pair p+1 needs that are not staged: the next tile's union
//     columns (for the copies of pair p+2) and entries (when the tile changes).
// Nothing issued in a pair is consumed in the same pair except T_{j-1} and v,
// which are issued before the copies, so their wait leaves the copies in flight.
// NW wavefronts per workgroup, UC staged rows per LDS buffer (SPW = UC / 4 NW
// staging loads per wavefront and chunk)
// PT: partial sums per tile (part[tile][R][2], for stacked jobs whose moments add
// up over their own tiles) instead of one per workgroup (part[workgroup][R][3]).
template <typename HT, typename XT, int VPL, int NCH, int GPW, int QS, bool DIRECT, int NW, int UC, bool PT>
__global__ __launch_bounds__(64 * NW, 1) void kpm_step_sell_kernel(SellStepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int C = 16 * VPL;
    constexpr int R = C * NCH;
    constexpr int SPW = UC / (4 * NW);
    static_assert(UC % (4 * NW) == 0, "LDS rows must split evenly over the wavefronts");
    __shared__ __align__(16) XT xs0[UC * C];
    __shared__ __align__(16) XT xs1[UC * C];
    // T_{j-1} of each wavefront's group rows, copied straight into LDS (lane l of
    // group j at slot j, bytes 16 l .. 16 l + 15: row s = l / 16, vectors of lane k = l % 16)
    constexpr bool TDMA = !DIRECT;
    __shared__ __align__(16) XT tp[TDMA ? NW * GPW * 64 * VPL : 1];
    // per-tile partials of each wavefront, two sets used by alternate tiles
    __shared__ double tred[PT ? 2 * NW * R * 2 : 1];
    static_assert(!(PT && DIRECT), "per-tile partials carry two sums");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: group loops branch on scalars
    const int s = lane >> 4;
    const int k = lane & 15;
    const int koff = k * VPL;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    int t_first, t_end, t_stride;
    if (a.sliced) {
        const int x = b & 7;
        t_first = (int)((int64_t)a.ntiles * x / 8) + (b >> 3);
        t_end = (int)((int64_t)a.ntiles * (x + 1) / 8);
        t_stride = G >> 3;
    } else {
        t_first = lb;
        t_end = a.ntiles;
        t_stride = G;
    }
    // the workgroup's tiles are t_first + k t_stride < t_end; on reversed steps
    // they are walked last to first, so a step begins with the rows the previous
    // step wrote last (still in L2 and the memory-side cache)
    if (a.reverse && t_first < t_end) {
        const int last = t_first + (t_end - 1 - t_first) / t_stride * t_stride;
        t_end = t_first - 1;                       // walk down to t_first
        t_first = last;
        t_stride = -t_stride;
    }
    // tile t of the walk exists (t_end is exclusive in the walk's direction)
    auto live = [&](int t) { return t_stride > 0 ? t < t_end : t > t_end; };

    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    const XT* prev = static_cast<const XT*>(a.prev);
    XT* out = static_cast<XT*>(a.out);
    const XT* __restrict__ vin = static_cast<const XT*>(a.v);
    const HT* __restrict__ ev = static_cast<const HT*>(a.ev);
    const uint16_t* __restrict__ eu = a.eu;
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;

    double d1[NCH][VPL], d2[NCH][VPL], d3[NCH][VPL];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            d1[c][v] = 0.0;
            d2[c][v] = 0.0;
            d3[c][v] = 0.0;
        }

    // entries, rows and lengths of this wavefront's groups of tile t
    auto load_groups = [&](int t, GroupRegs<HT, QS> (&gr)[GPW], int (&grow)[GPW], int (&glen)[GPW]) {
        const int g0 = ldu(a.tile_g0, t), g1 = ldu(a.tile_g0, t + 1);
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            const int g = g0 + wave + NW * j;
            const bool live = g < g1;
            const int L = live ? ldu(a.glen, g) : 0;
            const int64_t go = live ? ldu(a.goff, g) : 0;
            glen[j] = L;
            grow[j] = live ? a.grow[4 * g + s] : -1;
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                const int e = 16 * q + k;
                const bool ok = e < L;
                gr[j].eu[q] = ok ? (int)eu[go + 4 * e + s] : 0;
                if (ok) gr[j].ev[q] = ev[go + 4 * e + s];
                else zero(gr[j].ev[q]);
            }
        }
    };
    // union rows this lane copies for tile t (a valid dummy row past the union)
    auto load_ucols = [&](int t, int (&uc)[SPW]) {
        const int64_t u0 = ldu(a.tile_u0, t);
        const int U = (int)(ldu(a.tile_u0, t + 1) - u0);
#pragma unroll
        for (int m = 0; m < SPW; ++m) {
            const int u = 4 * (wave + NW * m) + s;
            uc[m] = a.ucols[u0 + (u < U ? u : 0)];
        }
    };
    auto stage = [&](XT* buf, int c, const int (&uc)[SPW]) {
#pragma unroll
        for (int m = 0; m < SPW; ++m) {
            const int i = wave + NW * m;
            const XT* src = cur + (size_t)uc[m] * R + c * C + koff;
            __builtin_amdgcn_global_load_lds(
                (const void*)src, (__attribute__((address_space(3))) void*)(buf + (size_t)4 * i * C), 16, 0, 0);
        }
    };

    GroupRegs<HT, QS> ec[GPW], en[GPW];
    int gc_row[GPW], gn_row[GPW], gc_len[GPW], gn_len[GPW];
    int ucs[SPW];
    int t = t_first, c = 0, p = 0;
    int r0 = 0, r0n = 0;
    if (live(t)) {
        load_groups(t, ec, gc_row, gc_len);
        r0 = ldu(a.tile_row0, t);
        load_ucols(t, ucs);
        stage(xs0, 0, ucs);
        if (NCH == 1) load_ucols(live(t + t_stride) ? t + t_stride : t, ucs);
    }

    int pend_tile = -1, pend_set = 0, tset = 0;
    // sum the wavefronts' partials of the tile finished in the previous pair
    auto emit_tile = [&]() {
        if (PT && pend_tile >= 0) {
            for (int i = threadIdx.x; i < 2 * R; i += 64 * NW) {
                double sum = tred[(pend_set * NW) * R * 2 + i];
#pragma unroll
                for (int w = 1; w < NW; ++w) sum += tred[(pend_set * NW + w) * R * 2 + i];
                a.part[(size_t)pend_tile * R * 2 + i] = sum;
            }
            pend_tile = -1;
        }
    };

    auto pair = [&](XT* buf, XT* nbuf) {
        wait_vm<0>();
        __syncthreads();
        emit_tile();
        if (c == 0 && p > 0) {
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                ec[j] = en[j];
                gc_row[j] = gn_row[j];
                gc_len[j] = gn_len[j];
            }
            r0 = r0n;
        }
        // T_{j-1} (and v) of this pair's rows.  Copied into LDS, T_{j-1} is waited for
        // with a count that leaves the staging copies issued after it in flight; a
        // register load would be waited for together with them.
        XT pv[GPW][VPL], vv[GPW][VPL];
        if constexpr (TDMA) {
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                const int row = gc_row[j];
                const XT* src = prev + (size_t)(row < 0 ? 0 : row) * R + c * C + koff;
                __builtin_amdgcn_global_load_lds((const void*)src,
                                                 (__attribute__((address_space(3))) void*)(tp + (wave * GPW + j) * 64 * VPL),
                                                 16, 0, 0);
            }
        } else {
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                const int row = gc_row[j];
                const size_t idx = (size_t)(row < 0 ? 0 : row) * R + c * C + koff;
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    if (beta && row >= 0) pv[j][v] = prev[idx + v];
                    else zero(pv[j][v]);
                    if (row >= 0) vv[j][v] = vin[idx + v];
                    else zero(vv[j][v]);
                }
            }
        }
        // copies of pair p+1 into the other buffer
        int t1 = t, c1 = c + 1;
        if (c1 == NCH) {
            c1 = 0;
            t1 = t + t_stride;
        }
        const bool has1 = live(t1);
        stage(nbuf, has1 ? c1 : 0, ucs);
        // products of every group of the pair
        AT acc[GPW][VPL];
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
#pragma unroll
            for (int v = 0; v < VPL; ++v) zero(acc[j][v]);
            int L = __builtin_amdgcn_readfirstlane(gc_len[j]);
            // opaque per pair: keeps the block conditions from being hoisted out of the loop
            asm volatile("" : "+s"(L));
            if (L > 0) {
                XT xa[4][VPL];
#pragma unroll
                for (int q = 0; q < QS; ++q) pin_values(ec[j].ev[q]);
                ld_block<0, HT, XT, VPL, QS>(xa, ec[j], buf, koff);
                run_blocks<0, HT, XT, AT, VPL, QS>(acc[j], ec[j], buf, koff, L, xa);
            }
        }
        if constexpr (TDMA) {
            // everything but the SPW staging copies issued after T_{j-1} has landed
            wait_vm<SPW>();
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                lds_ld<XT, VPL>(pv[j], tp + (wave * GPW + j) * 64 * VPL + lane * VPL);
                if (!beta || gc_row[j] < 0) {
#pragma unroll
                    for (int v = 0; v < VPL; ++v) zero(pv[j][v]);
                }
            }
        }
        // outputs and moment terms, then the stores
        XT o[GPW][VPL];
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            const int row = gc_row[j];
            XT cv[VPL];
            lds_ld<XT, VPL>(cv, buf + (row < 0 ? 0 : row - r0) * C + koff);
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                o[j][v] = comb(alpha, acc[j][v], pv[j][v]);
                const double t1v = row >= 0 ? nrm2(cv[v]) : 0.0;
                const double t2v = row >= 0 ? redot(o[j][v], cv[v]) : 0.0;
                const double t3v = DIRECT && row >= 0 ? redot(vv[j][v], cv[v]) : 0.0;
                // the chunk index is uniform: select this chunk's sums without indexing registers
#pragma unroll
                for (int cc = 0; cc < NCH; ++cc) {
                    if (cc == c) {
                        d1[cc][v] += t1v;
                        d2[cc][v] += t2v;
                        if (DIRECT) d3[cc][v] += t3v;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            const int row = gc_row[j];
            if (row >= 0) {
                const size_t idx = (size_t)row * R + c * C + koff;
#pragma unroll
                for (int v = 0; v < VPL; ++v) out[idx + v] = o[j][v];
            }
        }
        if (PT && c == NCH - 1) {
            // the tile's last chunk: fold its sums over the four DPP rows into this
            // wavefront's slot of the tile partials, summed over the wavefronts next pair
#pragma unroll
            for (int cc = 0; cc < NCH; ++cc)
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    d1[cc][v] += __shfl_xor(d1[cc][v], 16);
                    d2[cc][v] += __shfl_xor(d2[cc][v], 16);
                    d1[cc][v] += __shfl_xor(d1[cc][v], 32);
                    d2[cc][v] += __shfl_xor(d2[cc][v], 32);
                    if (s == 0) {
                        const int r = cc * C + koff + v;
                        tred[((tset * NW + wave) * R + r) * 2 + 0] = d1[cc][v];
                        tred[((tset * NW + wave) * R + r) * 2 + 1] = d2[cc][v];
                    }
                    d1[cc][v] = 0.0;
                    d2[cc][v] = 0.0;
                }
            pend_tile = t;
            pend_set = tset;
            tset ^= 1;
        }
        // loads for the next pairs: union columns of pair p+2's tile, the next tile's groups
        int t2 = t1, c2 = c1 + 1;
        if (c2 == NCH) {
            c2 = 0;
            t2 = t1 + t_stride;
        }
        if (c2 == 0) load_ucols(live(t2) ? t2 : t, ucs);
        if (c1 == 0 && has1) {
            load_groups(t1, en, gn_row, gn_len);
            r0n = ldu(a.tile_row0, t1);
        }
        t = t1;
        c = c1;
        ++p;
    };
    while (live(t)) {
        pair(xs0, xs1);
        if (!live(t)) break;
        pair(xs1, xs0);
    }
    wait_vm<0>();
    __syncthreads();
    if constexpr (PT) {
        emit_tile();
        return;
    }

    // fold the four DPP rows (same vectors), then the wavefronts
#pragma unroll
    for (int cc = 0; cc < NCH; ++cc)
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            d1[cc][v] += __shfl_xor(d1[cc][v], 16);
            d2[cc][v] += __shfl_xor(d2[cc][v], 16);
            d1[cc][v] += __shfl_xor(d1[cc][v], 32);
            d2[cc][v] += __shfl_xor(d2[cc][v], 32);
            if (DIRECT) {
                d3[cc][v] += __shfl_xor(d3[cc][v], 16);
                d3[cc][v] += __shfl_xor(d3[cc][v], 32);
            }
        }
    double* red = reinterpret_cast<double*>(xs0);   // [waves][R][3]
    if (s == 0) {
#pragma unroll
        for (int cc = 0; cc < NCH; ++cc)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                const int r = cc * C + koff + v;
                red[(wave * R + r) * 3 + 0] = d1[cc][v];
                red[(wave * R + r) * 3 + 1] = d2[cc][v];
                red[(wave * R + r) * 3 + 2] = DIRECT ? d3[cc][v] : 0.0;
            }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < R * 3; i += 64 * NW) {
        double sum = red[i];
#pragma unroll
        for (int w = 1; w < NW; ++w) sum += red[w * R * 3 + i];
        a.part[(size_t)lb * R * 3 + i] = sum;
    }
}

// The same step with the staging copies issued by a wavefront of their own.
//
// NWC compute wavefronts and one copy wavefront per workgroup.  Each wavefront
// has its own memory counters, so the compute wavefronts never have a direct
// global->LDS copy outstanding: with one outstanding, the compiler's wait
// insertion counts it as a possible LDS access of unknown order and turns every
// wait for an LDS read into a wait for all of them (lgkmcnt(0)), which drains
// the next block's operand reads before each block's products.  Without it,
// each block waits for its own four reads only.
//
// Pair p = (tile, chunk) of the walk, one barrier per pair:
//   copy wavefront:    barrier p; copies of pair p+1 into the buffer pair p-1
//                      used; union rows of the next tile; wait for all of it;
//   compute wavefront: barrier p; T_{j-1} (and v) of the pair's rows as register
//                      loads (T_{j-1} non-temporal); products from the buffer;
//                      outputs; the next tile's groups, then the stores of
//                      T_{j+1} (non-temporal).  The walk is unrolled over the
//                      chunks of a tile, so each pair's code has its chunk fixed.
// Both walk the same pairs, so they meet at the same barriers.  Tiles hold at
// most 4 GPW NWC rows (GPW groups of four rows per compute wavefront).
// PT: partial sums per tile (part[tile][R][2], stacked jobs) as in the kernel above.
template <typename HT, typename XT, int VPL, int NCH, int GPW, int QS, int NWC, int UC, bool PT>
__global__ __launch_bounds__(64 * (NWC + 1), 1) void kpm_step_sell_cw_kernel(SellStepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int C = 16 * VPL;
    constexpr int R = C * NCH;
    // a staged row chunk is RB = 256 bytes (16 complex FP64 or 32 complex FP32 vectors) or
    // 128 bytes (16 complex FP32 vectors); one 1 KiB copy moves RPI of them
    constexpr int RB = C * (int)sizeof(XT);
    static_assert(RB == 256 || RB == 128, "row chunks of 256 or 128 bytes");
    constexpr int LPR = RB / 16;        // lanes per staged row chunk (16 bytes each)
    constexpr int RPI = 64 / LPR;       // staged rows per copy instruction
    constexpr int NCP = UC / RPI;       // copies per pair, at most
    constexpr int NT = 64 * (NWC + 1);
    __shared__ __align__(16) XT xs0[UC * C];
    __shared__ __align__(16) XT xs1[UC * C];
    // per-tile partials of each compute wavefront, two sets used by alternate tiles
    __shared__ double tred[PT ? 2 * NWC * R * 2 : 1];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = lane >> 4;
    const int k = lane & 15;
    const int koff = k * VPL;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    int t_first, t_end, t_stride;
    if (a.sliced) {
        const int x = b & 7;
        t_first = (int)((int64_t)a.ntiles * x / 8) + (b >> 3);
        t_end = (int)((int64_t)a.ntiles * (x + 1) / 8);
        t_stride = G >> 3;
    } else {
        t_first = lb;
        t_end = a.ntiles;
        t_stride = G;
    }
    if (a.reverse && t_first < t_end) {
        const int last = t_first + (t_end - 1 - t_first) / t_stride * t_stride;
        t_end = t_first - 1;
        t_first = last;
        t_stride = -t_stride;
    }
    auto live = [&](int t) { return t_stride > 0 ? t < t_end : t > t_end; };
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);

    if (wave == NWC) {
        // ---- copy wavefront ----------------------------------------------------
        // byte offset of lane l's 16 bytes (piece l % LPR of union row RPI m + l / LPR)
        // within a chunk: the launcher checks n R sizeof(XT) < 2^32
        uint32_t uo[NCP], un[NCP];
        int U = 0, Un = 0;
        auto load_union = [&](int t, uint32_t (&dst)[NCP], int& cnt) {
            const int64_t u0 = ldu(a.tile_u0, t);
            cnt = (int)(ldu(a.tile_u0, t + 1) - u0);
#pragma unroll
            for (int m = 0; m < NCP; ++m) {
                const int u = RPI * m + lane / LPR;
                const int row = a.ucols[u0 + (u < cnt ? u : 0)];
                dst[m] = (uint32_t)row * (uint32_t)(R * sizeof(XT)) + (uint32_t)(16 * (lane % LPR));
            }
        };
        auto copies = [&](XT* buf, int c) {
            const char* base = reinterpret_cast<const char*>(cur + c * C);
            auto lbuf = (__attribute__((address_space(3))) XT*)buf;
#pragma unroll
            for (int m = 0; m < NCP; ++m) {
                if (RPI * m < U)
                    __builtin_amdgcn_global_load_lds((const void*)(base + uo[m]),
                                                     (__attribute__((address_space(3))) void*)(lbuf + RPI * m * C), 16,
                                                     0, 0);
            }
        };
        int t = t_first, c = 0;
        if (live(t)) {
            load_union(t, uo, U);
            copies(xs0, 0);
            if (live(t + t_stride)) load_union(t + t_stride, un, Un);
            wait_vm<0>();
        }
        auto pair = [&](XT* nbuf) {
            __syncthreads();
            int t1 = t, c1 = c + 1;
            if (c1 == NCH) {
                c1 = 0;
                t1 = t + t_stride;
            }
            if (live(t1)) {
                if (c1 == 0) {
#pragma unroll
                    for (int m = 0; m < NCP; ++m) uo[m] = un[m];
                    U = Un;
                }
                copies(nbuf, c1);
                if (c1 == 0 && live(t1 + t_stride)) load_union(t1 + t_stride, un, Un);
            }
            wait_vm<0>();
            t = t1;
            c = c1;
        };
        while (live(t)) {
            pair(xs1);
            if (!live(t)) break;
            pair(xs0);
        }
        __syncthreads();
        if constexpr (PT) return;
        __syncthreads();
    } else {
        // ---- compute wavefronts ------------------------------------------------
        const XT* prev = static_cast<const XT*>(a.prev);
        XT* out = static_cast<XT*>(a.out);
        const XT* __restrict__ vin = static_cast<const XT*>(a.v);
        const HT* __restrict__ ev = static_cast<const HT*>(a.ev);
        const uint16_t* __restrict__ eu = a.eu;
        const double alpha = a.alpha;

        double d1[NCH][VPL], d2[NCH][VPL], d3[NCH][VPL];
#pragma unroll
        for (int cc = 0; cc < NCH; ++cc)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                d1[cc][v] = 0.0;
                d2[cc][v] = 0.0;
                d3[cc][v] = 0.0;
            }
        auto load_groups = [&](int t, GroupRegs<HT, QS> (&gr)[GPW], int (&grow)[GPW], int (&glen)[GPW]) {
            const int g0 = ldu(a.tile_g0, t), g1 = ldu(a.tile_g0, t + 1);
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                const int g = g0 + wave + NWC * j;
                const bool lv = g < g1;
                const int L = lv ? ldu(a.glen, g) : 0;
                const int64_t go = lv ? ldu(a.goff, g) : 0;
                glen[j] = L;
                grow[j] = lv ? a.grow[4 * g + s] : -1;
#pragma unroll
                for (int q = 0; q < QS; ++q) {
                    const int e = 16 * q + k;
                    const bool ok = e < L;
                    gr[j].eu[q] = ok ? (int)eu[go + 4 * e + s] : 0;
                    if (ok) gr[j].ev[q] = ev[go + 4 * e + s];
                    else zero(gr[j].ev[q]);
                }
            }
        };
        GroupRegs<HT, QS> ec[GPW], en[GPW];
        int gc_row[GPW], gn_row[GPW], gc_len[GPW], gn_len[GPW];
        int t = t_first, p = 0;
        int r0 = 0, r0n = 0;
        if (live(t)) {
            load_groups(t, ec, gc_row, gc_len);
            r0 = ldu(a.tile_row0, t);
        }
        int pend_tile = -1, pend_set = 0, tset = 0;
        // sum the compute wavefronts' partials of the tile finished in the previous pair
        auto emit_tile = [&]() {
            if (PT && pend_tile >= 0) {
                for (int i = threadIdx.x; i < 2 * R; i += 64 * NWC) {
                    double sum = tred[(pend_set * NWC) * R * 2 + i];
#pragma unroll
                    for (int w = 1; w < NWC; ++w) sum += tred[(pend_set * NWC + w) * R * 2 + i];
                    a.part[(size_t)pend_tile * R * 2 + i] = sum;
                }
                pend_tile = -1;
            }
        };
        // the chunk index is a constant of each pair's code: the walk is unrolled over the
        // chunks of a tile, so the moment sums of chunk c are plain registers (a run-time
        // chunk index makes every sum an add into each chunk's register and a select)
        auto pair = [&](auto chunk, const XT* buf) {
            constexpr int c = decltype(chunk)::value;
            __syncthreads();
            emit_tile();
            // opaque per pair: keeps the compiler from specialising the whole walk on them
            int bflag = __builtin_amdgcn_readfirstlane(a.beta != 0 ? 1 : 0);
            int vflag = __builtin_amdgcn_readfirstlane(vin != nullptr ? 1 : 0);
            asm volatile("" : "+s"(bflag), "+s"(vflag));
            const bool beta = bflag != 0, has_v = vflag != 0;
            if (c == 0 && p > 0) {
#pragma unroll
                for (int j = 0; j < GPW; ++j) {
                    ec[j] = en[j];
                    gc_row[j] = gn_row[j];
                    gc_len[j] = gn_len[j];
                }
                r0 = r0n;
            }
            XT pv[GPW][VPL], vv[GPW][VPL];
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                const int row = gc_row[j];
                const size_t idx = (size_t)(row < 0 ? 0 : row) * R + c * C + koff;
                if (beta && row >= 0) {
                    ld_stream<XT, VPL>(pv[j], prev + idx);
                } else {
#pragma unroll
                    for (int v = 0; v < VPL; ++v) zero(pv[j][v]);
                }
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    if (has_v && row >= 0) vv[j][v] = vin[idx + v];
                    else zero(vv[j][v]);
                }
            }
            // Issue priority falls as a wavefront works through its groups (3, 2 from block 4
            // of the first group, 1, then 0 from block 4 of the second).  Two compute
            // wavefronts share each SIMD, and at equal priority the older one is issued
            // first, finishes its pair early and leaves the younger to run alone, its
            // latencies unhidden; ranking them by the work they have left keeps them
            // level (c1: 1.5 % per FP64 step, 2.5 % per FP32 step, same results).
            AT acc[GPW][VPL];
            static_assert(GPW == 1 || GPW == 2, "one or two groups per wavefront");
            auto group = [&](auto J) {
                constexpr int j = decltype(J)::value;
                __builtin_amdgcn_s_setprio(3 - 2 * j);
#pragma unroll
                for (int v = 0; v < VPL; ++v) zero(acc[j][v]);
                int L = __builtin_amdgcn_readfirstlane(gc_len[j]);
                asm volatile("" : "+s"(L));
                if (L > 0) {
                    XT xa[4][VPL];
#pragma unroll
                    for (int q = 0; q < QS; ++q) pin_values(ec[j].ev[q]);
                    ld_block<0, HT, XT, VPL, QS>(xa, ec[j], buf, koff);
                    run_blocks<0, HT, XT, AT, VPL, QS, 2 - 2 * j>(acc[j], ec[j], buf, koff, L, xa);
                }
            };
            group(std::integral_constant<int, 0>{});
            if constexpr (GPW == 2) group(std::integral_constant<int, 1>{});
            __builtin_amdgcn_s_setprio(0);
            XT o[GPW][VPL];
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                const int row = gc_row[j];
                XT cv[VPL];
                lds_ld<XT, VPL>(cv, buf + (row < 0 ? 0 : row - r0) * C + koff);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    o[j][v] = comb(alpha, acc[j][v], pv[j][v]);
                    // on a padded row (row < 0) the products, T_{j-1} and v are all zero (padded
                    // entries hold zero values), so o and vv are zero and only the norm needs the test
                    const double t1v = row >= 0 ? nrm2(cv[v]) : 0.0;
                    const double t2v = redot(o[j][v], cv[v]);
#pragma unroll
                    for (int cc = 0; cc < NCH; ++cc) {
                        if (cc == c) {
                            d1[cc][v] += t1v;
                            d2[cc][v] += t2v;
                        }
                    }
                }
                if (has_v) {
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        const double t3v = redot(vv[j][v], cv[v]);
#pragma unroll
                        for (int cc = 0; cc < NCH; ++cc)
                            if (cc == c) d3[cc][v] += t3v;
                    }
                }
            }
            if (PT && c == NCH - 1) {
                // the tile's last chunk: fold its sums over the four DPP rows into this
                // wavefront's slot of the tile partials, summed over the wavefronts next pair
#pragma unroll
                for (int cc = 0; cc < NCH; ++cc)
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        d1[cc][v] += __shfl_xor(d1[cc][v], 16);
                        d2[cc][v] += __shfl_xor(d2[cc][v], 16);
                        d1[cc][v] += __shfl_xor(d1[cc][v], 32);
                        d2[cc][v] += __shfl_xor(d2[cc][v], 32);
                        if (s == 0) {
                            const int r = cc * C + koff + v;
                            tred[((tset * NWC + wave) * R + r) * 2 + 0] = d1[cc][v];
                            tred[((tset * NWC + wave) * R + r) * 2 + 1] = d2[cc][v];
                        }
                        d1[cc][v] = 0.0;
                        d2[cc][v] = 0.0;
                    }
                pend_tile = t;
                pend_set = tset;
                tset ^= 1;
            }
            const int t1 = c + 1 == NCH ? t + t_stride : t;
            // the next tile's groups before this pair's stores: waiting for them then
            // leaves the stores in flight
            if (c + 1 == NCH && live(t1)) {
                load_groups(t1, en, gn_row, gn_len);
                r0n = ldu(a.tile_row0, t1);
            }
#pragma unroll
            for (int j = 0; j < GPW; ++j) {
                const int row = gc_row[j];
                if (row >= 0) {
                    const size_t idx = (size_t)row * R + c * C + koff;
                    st_stream<XT, VPL>(out + idx, o[j]);
                }
            }
            t = t1;
            ++p;
        };
        // pairs alternate between the two buffers: with an even chunk count chunk c always
        // uses buffer c % 2, with one chunk the tiles alternate
        static_assert(NCH == 1 || NCH == 2 || NCH == 4, "one, two or four chunks");
        while (live(t)) {
            if constexpr (NCH == 1) {
                pair(std::integral_constant<int, 0>{}, xs0);
                if (!live(t)) break;
                pair(std::integral_constant<int, 0>{}, xs1);
            } else {
                pair(std::integral_constant<int, 0>{}, xs0);
                pair(std::integral_constant<int, 1>{}, xs1);
                if constexpr (NCH == 4) {
                    pair(std::integral_constant<int, 2>{}, xs0);
                    pair(std::integral_constant<int, 3>{}, xs1);
                }
            }
        }
        // every wavefront is past its last read of the buffers: the partials go into xs0
        __syncthreads();
        if constexpr (PT) {
            emit_tile();
            return;
        }
#pragma unroll
        for (int cc = 0; cc < NCH; ++cc)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                d1[cc][v] += __shfl_xor(d1[cc][v], 16);
                d2[cc][v] += __shfl_xor(d2[cc][v], 16);
                d3[cc][v] += __shfl_xor(d3[cc][v], 16);
                d1[cc][v] += __shfl_xor(d1[cc][v], 32);
                d2[cc][v] += __shfl_xor(d2[cc][v], 32);
                d3[cc][v] += __shfl_xor(d3[cc][v], 32);
            }
        double* red = reinterpret_cast<double*>(xs0);   // [compute waves][R][3]
        if (s == 0) {
#pragma unroll
            for (int cc = 0; cc < NCH; ++cc)
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    const int r = cc * C + koff + v;
                    red[(wave * R + r) * 3 + 0] = d1[cc][v];
                    red[(wave * R + r) * 3 + 1] = d2[cc][v];
                    red[(wave * R + r) * 3 + 2] = d3[cc][v];
                }
        }
        __syncthreads();
    }
    const double* red = reinterpret_cast<const double*>(xs0);
    for (int i = threadIdx.x; i < R * 3; i += NT) {
        double sum = red[i];
#pragma unroll
        for (int w = 1; w < NWC; ++w) sum += red[w * R * 3 + i];
        a.part[(size_t)lb * R * 3 + i] = sum;
    }
}

template <typename HT, typename XT, int VPL, int NCH, int GPW, int QS, int NW>
hipError_t launch_sell(const SellStepArgs& a, hipStream_t s)
{
    if (a.copy_wave) {
        // one copy wavefront beside kSellCwWaves compute wavefronts; 32-bit copy offsets
        constexpr int NWC = kSellCwWaves;
        if (a.tile_rows > 4 * GPW * NWC || a.ucap > kSellUcap ||
            (uint64_t)a.n * (uint64_t)a.R * sizeof(XT) >= (uint64_t(1) << 32) || (a.tile_part && a.v))
            return hipErrorInvalidValue;
        if (a.tile_part)
            hipLaunchKernelGGL((kpm_step_sell_cw_kernel<HT, XT, VPL, NCH, GPW, QS, NWC, kSellUcap, true>), dim3(a.grid),
                               dim3(64 * (NWC + 1)), 0, s, a);
        else
            hipLaunchKernelGGL((kpm_step_sell_cw_kernel<HT, XT, VPL, NCH, GPW, QS, NWC, kSellUcap, false>),
                               dim3(a.grid), dim3(64 * (NWC + 1)), 0, s, a);
        return hipGetLastError();
    }
    constexpr int UC = kSellUcap, UCT = kSellUcapTiled;
    if (a.tile_part) {
        if (a.v || a.ucap > UCT) return hipErrorInvalidValue;
        hipLaunchKernelGGL((kpm_step_sell_kernel<HT, XT, VPL, NCH, GPW, QS, false, NW, UCT, true>), dim3(a.grid),
                           dim3(64 * NW), 0, s, a);
    } else if (a.v) {
        hipLaunchKernelGGL((kpm_step_sell_kernel<HT, XT, VPL, NCH, GPW, QS, true, NW, UC, false>), dim3(a.grid),
                           dim3(64 * NW), 0, s, a);
    } else {
        hipLaunchKernelGGL((kpm_step_sell_kernel<HT, XT, VPL, NCH, GPW, QS, false, NW, UC, false>), dim3(a.grid),
                           dim3(64 * NW), 0, s, a);
    }
    return hipGetLastError();
}

template <typename HT, typename XT, int VPL, int GPW, int QS>
hipError_t launch_sell_r(const SellStepArgs& a, hipStream_t s)
{
    constexpr int NW = kSellWaves;
    switch (a.R / (16 * VPL)) {
    case 1: return launch_sell<HT, XT, VPL, 1, GPW, QS, NW>(a, s);
    case 2: return launch_sell<HT, XT, VPL, 2, GPW, QS, NW>(a, s);
    case 4:
        if constexpr (VPL == 1) {
            if (a.tile_part) return hipErrorInvalidValue;   // stacked jobs carry 16 or 32 vectors
            return launch_sell<HT, XT, VPL, 4, GPW, QS, NW>(a, s);
        }
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

// Two groups of four rows per wavefront (tiles of at most 64 rows): with the
// next tile's entries double-buffered in registers, four groups of complex
// FP64 entries per wavefront exceed the 256 registers a lane has at two
// wavefronts per SIMD.  (Four groups of complex FP32 entries fit, and were
// measured slower on the 1.05 deg 8 x 8 cell: 580 against 558 us per step.)
template <typename HT, typename XT, int VPL>
hipError_t launch_sell_shape(const SellStepArgs& a, hipStream_t s)
{
    if (a.max_len > 48) return launch_sell_r<HT, XT, VPL, 2, 4>(a, s);
    return launch_sell_r<HT, XT, VPL, 2, 3>(a, s);
}

// 16 complex FP32 vectors: 128-byte row chunks (eight staged rows per copy), the
// copy-wavefront kernel only
template <typename HT>
hipError_t launch_sell_f32x16(const SellStepArgs& a, hipStream_t s)
{
    constexpr int NWC = kSellCwWaves, GPW = 2, UC = kSellUcap;
    if (!a.copy_wave || a.tile_rows > 4 * GPW * NWC || a.ucap > UC || (a.tile_part && a.v) ||
        (uint64_t)a.n * 16 * sizeof(float2) >= (uint64_t(1) << 32))
        return hipErrorInvalidValue;
    const dim3 grid(a.grid), block(64 * (NWC + 1));
    if (a.max_len > 48) {
        if (a.tile_part)
            hipLaunchKernelGGL((kpm_step_sell_cw_kernel<HT, float2, 1, 1, GPW, 4, NWC, UC, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((kpm_step_sell_cw_kernel<HT, float2, 1, 1, GPW, 4, NWC, UC, false>), grid, block, 0, s, a);
    } else {
        if (a.tile_part)
            hipLaunchKernelGGL((kpm_step_sell_cw_kernel<HT, float2, 1, 1, GPW, 3, NWC, UC, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((kpm_step_sell_cw_kernel<HT, float2, 1, 1, GPW, 3, NWC, UC, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

// dst[i] = src[map[i]] (0 where map[i] < 0): the SELL values from the CSR values
template <typename E>
__global__ void sell_fill_kernel(const int64_t* __restrict__ map, const E* __restrict__ src, E* __restrict__ dst,
                                 int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = map[i];
        E e;
        if (m >= 0) e = src[m];
        else zero(e);
        dst[i] = e;
    }
}

}  // namespace

bool sell_supported(int mat_type, int vec_type, int R, bool copy_wave)
{
    const bool cpx64 = vec_type == VEC_CPX_F64 && (mat_type == MAT_CPX_F64 || mat_type == MAT_REAL_F64);
    const bool cpx32 = vec_type == VEC_CPX_F32 && (mat_type == MAT_CPX_F32 || mat_type == MAT_REAL_F32);
    const bool real64 = vec_type == VEC_REAL_F64 && mat_type == MAT_REAL_F64;
    if (cpx64) return R == 16 || R == 32 || R == 64;
    if (cpx32) return R == 32 || R == 64 || (copy_wave && R == 16);
    if (real64) return R == 32 || R == 64;   // two FP64 vectors per lane, 32 per 256-byte chunk
    return false;
}

hipError_t launch_kpm_step_sell(int mat_type, int vec_type, const SellStepArgs& a, hipStream_t s)
{
    if (a.grid % 8 != 0 || a.grid <= 0 || !sell_supported(mat_type, vec_type, a.R, a.copy_wave != 0) ||
        a.tile_rows > kSellMaxRows || a.max_len > kSellMaxLen)
        return hipErrorInvalidValue;
    if (mat_type == MAT_CPX_F64) return launch_sell_shape<double2, double2, 1>(a, s);
    if (mat_type == MAT_REAL_F64 && vec_type == VEC_REAL_F64) return launch_sell_shape<double, double, 2>(a, s);
    if (mat_type == MAT_REAL_F64) return launch_sell_shape<double, double2, 1>(a, s);
    if (mat_type == MAT_CPX_F32)
        return a.R == 16 ? launch_sell_f32x16<float2>(a, s) : launch_sell_shape<float2, float2, 2>(a, s);
    if (mat_type == MAT_REAL_F32)
        return a.R == 16 ? launch_sell_f32x16<float>(a, s) : launch_sell_shape<float, float2, 2>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_sell_fill(const int64_t* map, const void* src, void* dst, int elem_bytes, int64_t count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int blocks = (int)std::min<int64_t>(4096, (count + 255) / 256);
    switch (elem_bytes) {
    case 4:
        hipLaunchKernelGGL(sell_fill_kernel<float>, dim3(blocks), dim3(256), 0, s, map, (const float*)src, (float*)dst,
                           count);
        break;
    case 8:
        hipLaunchKernelGGL(sell_fill_kernel<double>, dim3(blocks), dim3(256), 0, s, map, (const double*)src,
                           (double*)dst, count);
        break;
    case 16:
        hipLaunchKernelGGL(sell_fill_kernel<double2>, dim3(blocks), dim3(256), 0, s, map, (const double2*)src,
                           (double2*)dst, count);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace leo2d
