// dgemm.hip -- FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// Used by the KPM conductivity: the moment matrix C_b = D_b^T B_b of every
// target (M x M outputs, n-long reductions; locality.cpp:6660-6700 forms the
// same numbers one dot product at a time) and the two GEMMs of the 2D
// Jackson/DCT-III transform (param_tools.cpp:838-905).
//
//   C[b][s](m, n) = sum_{k in split s} A_b(m, k) B_b(k, n)
//   A_b(m, k) = A[b sab + m sam + k sak],  B_b(k, n) = B[b sbb + n sbn + k sbk],
//   C[b][s](m, n) = C[b scb + s scs + m scm + n scn]
//
// Split s covers k in [s kc, min(K, (s+1) kc)), kc = ceil(K / S) rounded up to
// the 16-deep k tile, so the partial products of one (b, m, n) are summed by
// the caller in split order: results do not depend on the launch shape.
//
// Tiles: 128 x 128 outputs per 256-thread workgroup, four wavefronts of 64 x 64
// (4 x 4 MFMA blocks of 16 x 16, 64 FP64 accumulators per lane), k in tiles of
// 16 staged through LDS in two buffers: the next tile's global loads are in
// flight while the current tile's 64 MFMAs per wavefront run, one barrier per
// tile.  LDS holds each operand tile as [row][16 + 2]: a 16-lane MFMA operand
// read (16 rows, one k) then touches 16 distinct 4-bank groups.
// Operands whose k stride is 1 are loaded 16 bytes at a time along k (the
// conductivity's D and B sets, [m][n] with n contiguous); others one element
// at a time along the row index (the transform's column-major matrices).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace leo2d {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

constexpr int kTile = 128;      // outputs per workgroup along m and n
constexpr int kDepth = 16;      // k per LDS tile
constexpr int kPitch = 18;      // LDS row pitch in doubles (16 + 2)
constexpr int kThreads = 256;

// one operand tile (128 rows x 16 k) from global memory into registers:
// KC (k contiguous, 16-byte aligned rows): thread t loads chunks q*256 + t,
// chunk c = (row c / 8, k pair 2 (c % 8)); otherwise element e = q*256 + t is
// (row e % 128, k e / 128)
template <bool KC>
struct TileRegs {
    d2v v[KC ? 4 : 0];
    double s[KC ? 0 : 8];
};

template <bool KC>
__device__ inline void load_tile(TileRegs<KC>& r, const double* __restrict__ P, int64_t srow, int64_t sk, int row0,
                                 int rows, int k0, int kend, int t)
{
    if constexpr (KC) {
        if (row0 + kTile <= rows && k0 + kDepth <= kend) {
            // interior tile (uniform branch): unconditional 16-byte loads
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = q * kThreads + t;
                r.v[q] = *reinterpret_cast<const d2v*>(P + (int64_t)(row0 + (c >> 3)) * srow + k0 + 2 * (c & 7));
            }
        } else {
            // edge tile: element loads from clamped addresses, zero outside
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = q * kThreads + t;
                const int row = row0 + (c >> 3), k = k0 + 2 * (c & 7);
                const bool rok = row < rows;
                const bool v0 = rok && k < kend, v1 = rok && k + 1 < kend;
                const int64_t base = (int64_t)row * srow + k;
                const double x0 = P[v0 ? base : 0];
                const double x1 = P[v1 ? base + 1 : 0];
                r.v[q] = d2v{v0 ? x0 : 0.0, v1 ? x1 : 0.0};
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = q * kThreads + t;
            const int row = row0 + (e & (kTile - 1)), k = k0 + (e >> 7);
            const bool ok = row < rows && k < kend;
            const double x = P[ok ? (int64_t)row * srow + (int64_t)k * sk : 0];
            r.s[q] = ok ? x : 0.0;
        }
    }
}

template <bool KC>
__device__ inline void store_tile(const TileRegs<KC>& r, double* __restrict__ L, int t)
{
    if constexpr (KC) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = q * kThreads + t;
            *reinterpret_cast<d2v*>(L + (c >> 3) * kPitch + 2 * (c & 7)) = r.v[q];
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = q * kThreads + t;
            L[(e & (kTile - 1)) * kPitch + (e >> 7)] = r.s[q];
        }
    }
}

template <bool KCA, bool KCB>
__global__ __launch_bounds__(kThreads, 2) void dgemm_kernel(DgemmArgs a)
{
    __shared__ double la[2][kTile * kPitch];
    __shared__ double lb[2][kTile * kPitch];

    // logical block: consecutive ids on one XCD (dispatch deals workgroups to
    // the 8 XCDs in turn), so the output tiles of one (batch, split) -- which
    // read the same k range of A and B -- share that XCD's L2
    const int G = gridDim.x;
    const int bid = blockIdx.x;
    const int lid = (G % 8 == 0) ? (bid % 8) * (G / 8) + bid / 8 : bid;
    const int tm = (a.M + kTile - 1) / kTile, tn = (a.N + kTile - 1) / kTile;
    const int tile = lid % (tm * tn);
    const int s = (lid / (tm * tn)) % a.splits;
    const int b = lid / (tm * tn * a.splits);
    const int m0 = (tile % tm) * kTile, n0 = (tile / tm) * kTile;
    const int kbeg = s * a.kchunk;
    const int kend = min(a.K, kbeg + a.kchunk);
    const double* A = a.A + (int64_t)b * a.sab;
    const double* B = a.B + (int64_t)b * a.sbb;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int wm = (w & 1) * 64, wn = (w >> 1) * 64;   // the wavefront's 64 x 64 quarter
    const int li = lane & 15, lk = lane >> 4;

    d4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    TileRegs<KCA> ra;
    TileRegs<KCB> rb;
    const int ntiles = kend > kbeg ? (kend - kbeg + kDepth - 1) / kDepth : 0;
    if (ntiles > 0) {
        load_tile<KCA>(ra, A, a.sam, a.sak, m0, a.M, kbeg, kend, t);
        load_tile<KCB>(rb, B, a.sbn, a.sbk, n0, a.N, kbeg, kend, t);
        store_tile<KCA>(ra, la[0], t);
        store_tile<KCB>(rb, lb[0], t);
    }
    __syncthreads();
    for (int kt = 0; kt < ntiles; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < ntiles;
        if (more) {
            const int k0 = kbeg + (kt + 1) * kDepth;
            load_tile<KCA>(ra, A, a.sam, a.sak, m0, a.M, k0, kend, t);
            load_tile<KCB>(rb, B, a.sbn, a.sbk, n0, a.N, k0, kend, t);
        }
        const double* pa = la[cur] + (wm + li) * kPitch + lk;
        const double* pb = lb[cur] + (wn + li) * kPitch + lk;
#pragma unroll
        for (int ks = 0; ks < kDepth / 4; ++ks) {
            double fa[4], fb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) fa[i] = pa[i * 16 * kPitch + 4 * ks];
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[j] = pb[j * 16 * kPitch + 4 * ks];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            store_tile<KCA>(ra, la[cur ^ 1], t);
            store_tile<KCB>(rb, lb[cur ^ 1], t);
        }
        __syncthreads();
    }

    // result layout of v_mfma_f64_16x16x4f64: lane (li, lk) of block (i, j) holds
    // rows lk + 4 r (r = 0..3) of column li (rows interleaved across the lane groups)
    double* C = a.C + (int64_t)b * a.scb + (int64_t)s * a.scs;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn + j * 16 + li;
            if (n >= a.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + i * 16 + lk + 4 * r;
                if (m < a.M) C[(int64_t)m * a.scm + (int64_t)n * a.scn] = acc[i][j][r];
            }
        }
}

}  // namespace

hipError_t launch_dgemm(const DgemmArgs& in, hipStream_t s)
{
    DgemmArgs a = in;
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return hipSuccess;
    if (a.K < 0 || a.splits < 1) return hipErrorInvalidValue;
    a.kchunk = (a.K + a.splits - 1) / a.splits;
    a.kchunk = (a.kchunk + kDepth - 1) / kDepth * kDepth;
    if (a.kchunk == 0) a.kchunk = kDepth;
    // 16-byte loads along k: unit k stride, even row stride, 16-byte aligned base
    const bool kca = a.sak == 1 && a.sam % 2 == 0 && a.sab % 2 == 0 && (reinterpret_cast<uintptr_t>(a.A) & 15) == 0;
    const bool kcb = a.sbk == 1 && a.sbn % 2 == 0 && a.sbb % 2 == 0 && (reinterpret_cast<uintptr_t>(a.B) & 15) == 0;
    const int64_t tiles = (int64_t)((a.M + kTile - 1) / kTile) * ((a.N + kTile - 1) / kTile);
    const int64_t blocks = tiles * a.splits * a.batch;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    const dim3 g((unsigned)blocks);
    if (kca && kcb) hipLaunchKernelGGL((dgemm_kernel<true, true>), g, dim3(kThreads), 0, s, a);
    else if (kca) hipLaunchKernelGGL((dgemm_kernel<true, false>), g, dim3(kThreads), 0, s, a);
    else if (kcb) hipLaunchKernelGGL((dgemm_kernel<false, true>), g, dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL((dgemm_kernel<false, false>), g, dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace leo2d
