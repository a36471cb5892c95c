// kpm.hip -- Chebyshev (KPM) recursion over a block of R start vectors.
//
// One launch = one Chebyshev step for all R vectors:
//     T_{j+1} = 2 H T_j - T_{j-1}        (T_1 = H T_0 on the first step)
// fused with the per-vector partial sums of <T_j|T_j> and Re<T_{j+1}|T_j>
// (and Re<v|T_j> for the reference's vector-start scheme).  Replaces the
// reference's per-vector SpMV loop (locality.cpp:5945-6346 calling
// SpMatrix::vectorMultiply, spmatrix.cpp:270-411): H is streamed once per step
// for all R vectors instead of once per vector, and the two vector copies and
// the subtract of the reference loop are fused away (T_{j+1} overwrites T_{j-1}).
//
// Layout: vectors row-major [n][R] so that the gather of one column index
// reads R contiguous entries (R = 64 complex FP64 = 1 KiB per nonzero per
// wavefront, one fully coalesced load).  Lanes of a wavefront map to vectors;
// a wavefront owns 64/R rows when R < 64.
//
// Scheduling: a fixed grid of G workgroups (G multiple of 8) sweeps the rows
// in strides of G * rows-per-workgroup.  Hardware dispatches workgroup b to
// XCD (b mod 8); the logical slot (b mod 8) * G/8 + b/8 gives every XCD one
// contiguous block of G/8 workgroup tiles per sweep, so at any moment each
// XCD works on a compact window of rows whose gathered neighbours (T_j rows)
// stay in its L2.  With the spatial (Morton) row order applied by the host,
// that window is a compact 2D patch of the lattice.
//
// Reductions are deterministic: fixed per-lane order, fixed shuffle tree,
// fixed cross-wavefront order, one partial per workgroup and vector, then a
// fixed-order tree over workgroups in kpm_reduce.  No floating-point atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernels.h"
#include "kpm_common.h"

namespace leo2d {
namespace {

// ---- the fused step --------------------------------------------------------
// LPR: lanes per row (= min(R, 64)); VPL: vectors per lane (R / 64 when R > 64).
template <typename HT, typename XT, int LPR, int VPL>
__global__ __launch_bounds__(kThreads) void kpm_step_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int RPW = 64 / LPR;          // rows per wavefront
    constexpr int RPI = RPW * kWaves;      // rows per workgroup iteration
    const int R = LPR * VPL;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sub = lane / LPR;
    const int r0 = lane % LPR;

    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);   // XCD-contiguous slot in each sweep
    const int n = a.n;
    const int row_end = n;
    const int stride = G * RPI;

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const int32_t* __restrict__ col = a.col;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    const XT* prev = static_cast<const XT*>(a.prev);
    XT* out = static_cast<XT*>(a.out);
    const XT* __restrict__ vin = static_cast<const XT*>(a.v);
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;
    const bool direct = vin != nullptr;

    double d1[VPL], d2[VPL], d3[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) { d1[v] = 0.0; d2[v] = 0.0; d3[v] = 0.0; }

    for (int row0 = lb * RPI + wave * RPW; row0 < row_end; row0 += stride) {
        int row = row0 + sub;
        if (RPW == 1) row = __builtin_amdgcn_readfirstlane(row);
        if (row >= row_end) continue;
        AT acc[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) zero(acc[v]);
        int64_t q = rp[row];
        const int64_t e = rp[row + 1];
        for (; q + 4 <= e; q += 4) {
            const int c0 = col[q], c1 = col[q + 1], c2 = col[q + 2], c3 = col[q + 3];
            const HT h0 = val[q], h1 = val[q + 1], h2 = val[q + 2], h3 = val[q + 3];
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                const int r = r0 + v * LPR;
                const XT x0 = cur[(size_t)c0 * R + r];
                const XT x1 = cur[(size_t)c1 * R + r];
                const XT x2 = cur[(size_t)c2 * R + r];
                const XT x3 = cur[(size_t)c3 * R + r];
                mac(acc[v], h0, x0);
                mac(acc[v], h1, x1);
                mac(acc[v], h2, x2);
                mac(acc[v], h3, x3);
            }
        }
        for (; q < e; ++q) {
            const int c0 = col[q];
            const HT h0 = val[q];
#pragma unroll
            for (int v = 0; v < VPL; ++v) mac(acc[v], h0, cur[(size_t)c0 * R + r0 + v * LPR]);
        }
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            const size_t idx = (size_t)row * R + r0 + v * LPR;
            const XT c = cur[idx];
            XT p;
            if (beta) p = prev[idx];
            else zero(p);
            const XT o = comb(alpha, acc[v], p);
            out[idx] = o;
            d1[v] += nrm2(c);
            d2[v] += redot(o, c);
            if (direct) d3[v] += redot(vin[idx], c);
        }
    }

    // fold the row groups of a wavefront (lanes r0, r0+LPR, ...)
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1) {
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            d1[v] += __shfl_xor(d1[v], off);
            d2[v] += __shfl_xor(d2[v], off);
            d3[v] += __shfl_xor(d3[v], off);
        }
    }
    __shared__ double red[kWaves][LPR * VPL][3];
    if (sub == 0) {
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            red[wave][r0 + v * LPR][0] = d1[v];
            red[wave][r0 + v * LPR][1] = d2[v];
            red[wave][r0 + v * LPR][2] = d3[v];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        const int r = t / 3, k = t % 3;
        double s = red[0][r][k];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += red[w][r][k];
        a.part[((size_t)lb * R + r) * 3 + k] = s;
    }
}

// ---- wavefront-per-row step for R >= 64 --------------------------------------
// A row's (column, value) entries are fetched with one coalesced load per 64
// entries (lane k holds entry k), then broadcast lane by lane with readlane
// into scalar registers: the gather address becomes scalar base + lane offset
// and BATCH gathers are issued back to back with no dependent index load in
// between.  The next row's entries are prefetched while the current row's
// gathers are in flight.
__device__ inline int rl(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ inline double rl(double v, int k)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float rl(float v, int k) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k)); }
__device__ inline double2 rl(double2 v, int k) { return make_double2(rl(v.x, k), rl(v.y, k)); }
__device__ inline float2 rl(float2 v, int k) { return make_float2(rl(v.x, k), rl(v.y, k)); }

__device__ inline int64_t uniform64(int64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v & 0xffffffffLL));
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <typename HT>
__device__ inline HT zero_h()
{
    HT z;
    zero(z);
    return z;
}

template <typename HT, typename XT, int VPL, int BATCH>
__global__ __launch_bounds__(kThreads) void kpm_step_wave_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int NW = kWaves;
    const int R = 64 * VPL;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    // rows of this workgroup: interleaved over the whole matrix, or (xcd_rows
    // > 0) a grid-stride walk over its XCD's own contiguous slice of rows
    int first, n, stride;
    if (a.xcd_rows > 0) {
        first = (b & 7) * a.xcd_rows + (b >> 3) * NW;
        n = min(a.n, ((b & 7) + 1) * a.xcd_rows);
        stride = (G >> 3) * NW;
    } else {
        first = lb * NW;
        n = a.n;
        stride = G * NW;
    }

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const int32_t* __restrict__ col = a.col;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    const XT* prev = static_cast<const XT*>(a.prev);
    XT* out = static_cast<XT*>(a.out);
    const XT* __restrict__ vin = static_cast<const XT*>(a.v);
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;
    const bool direct = vin != nullptr;

    double d1[VPL], d2[VPL], d3[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) { d1[v] = 0.0; d2[v] = 0.0; d3[v] = 0.0; }

    int row = __builtin_amdgcn_readfirstlane(first + wave);
    // prefetch the first row's entries
    int64_t q0 = row < n ? rp[row] : 0, q1 = row < n ? rp[row + 1] : 0;
    q0 = uniform64(q0);
    q1 = uniform64(q1);
    int my_c = (row < n && q0 + lane < q1) ? *(col + q0 + lane) : 0;
    HT my_h = (row < n && q0 + lane < q1) ? *(val + q0 + lane) : zero_h<HT>();

    for (; row < n; row += stride) {
        const int next = row + stride;
        int64_t nq0 = 0, nq1 = 0;
        if (next < n) {
            nq0 = rp[next];
            nq1 = rp[next + 1];
        }
        AT acc[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) zero(acc[v]);
        int64_t base = q0;
        int c_reg = my_c;
        HT h_reg = my_h;
        for (;;) {
            const int cnt = (int)min((int64_t)64, q1 - base);
            for (int k = 0; k < cnt; k += BATCH) {
                XT x[BATCH][VPL];
#pragma unroll
                for (int u = 0; u < BATCH; ++u) {
                    const int kk = min(k + u, cnt - 1);
                    const int c = rl(c_reg, kk);
#pragma unroll
                    for (int v = 0; v < VPL; ++v) x[u][v] = cur[(size_t)c * R + lane + 64 * v];
                }
#pragma unroll
                for (int u = 0; u < BATCH; ++u) {
                    if (k + u < cnt) {
                        const HT h = rl(h_reg, k + u);
#pragma unroll
                        for (int v = 0; v < VPL; ++v) mac(acc[v], h, x[u][v]);
                    }
                }
            }
            base += 64;
            if (base >= q1) break;
            c_reg = (base + lane < q1) ? *(col + base + lane) : 0;
            h_reg = (base + lane < q1) ? *(val + base + lane) : zero_h<HT>();
        }
        // next row's entries (issued before this row's epilogue traffic)
        nq0 = uniform64(nq0);
        nq1 = uniform64(nq1);
        if (next < n) {
            my_c = (nq0 + lane < nq1) ? *(col + nq0 + lane) : 0;
            my_h = (nq0 + lane < nq1) ? *(val + nq0 + lane) : zero_h<HT>();
        }
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            const size_t idx = (size_t)row * R + lane + 64 * v;
            const XT c = cur[idx];
            XT p;
            if (beta) p = *(prev + idx);
            else zero(p);
            const XT o = comb(alpha, acc[v], p);
            out[idx] = o;
            d1[v] += nrm2(c);
            d2[v] += redot(o, c);
            if (direct) d3[v] += redot(vin[idx], c);
        }
        q0 = nq0;
        q1 = nq1;
    }

    __shared__ double red[NW][64 * VPL][3];
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        red[wave][lane + 64 * v][0] = d1[v];
        red[wave][lane + 64 * v][1] = d2[v];
        red[wave][lane + 64 * v][2] = d3[v];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        const int r = t / 3, k = t % 3;
        double s = red[0][r][k];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += red[w][r][k];
        a.part[((size_t)lb * R + r) * 3 + k] = s;
    }
}

template <int B> struct alignas(4) ColChunk { int c[B]; };
template <typename HT, int B> struct ValChunk { HT h[B]; };

// ---- variant 14: vector chunks split over the XCDs -----------------------------
// The gathers of the wavefront-per-row kernels are served by L2 about 90 % of
// the time; the rest (neighbours far apart in the row order, such as across
// the periodic boundary) come from the memory-side cache or HBM and dominate
// the step.  Each XCD's L2 holds rows of T_j as 1 KiB (R = 64, complex FP64)
// units.  This kernel gives each XCD a chunk of 16 vectors only: XCD x works on
// chunk x mod (R/16) and, when R/16 < 8, on part x / (R/16) of the rows.  An
// XCD's L2 then holds 256-byte row chunks, so it keeps R/16 times more
// neighbour rows resident; every XCD reads the matrix entries of its rows
// (R/16 = 4 parts' worth of H overall at R = 64).
// Lane = (row slot 0..3) x (vector 0..15): the 16 lanes of a slot load 16 of
// their row's (column, value) entries in one access and pass them to each
// other with DPP row broadcasts; gathers for 16 entries are in flight at once
// (lanes past their row's end are masked off), and the next chunk of entries
// is loaded before the current one is consumed.  A workgroup's 4 wavefronts
// own 16 consecutive rows; the XCD's workgroups sweep the XCD's rows in that
// interleave, so the XCD works on one compact window of rows at a time.
template <typename HT, typename XT, int NCH>
__global__ __launch_bounds__(kThreads) void kpm_step_xchunk_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int C = 16;                 // vectors per chunk
    constexpr int R = C * NCH;
    constexpr int P = 8 / NCH;            // row parts
    constexpr int RB = kWaves * 4;        // rows per workgroup block
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane >> 4;
    const int vi = lane & 15;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int xcd = b & 7;
    const int GW = G >> 3;
    const int w = b >> 3;
    const int lb = xcd * GW + w;
    const int chunk = xcd % NCH;
    const int part = xcd / NCH;
    const int lo = (int)((int64_t)a.n * part / P);
    const int hi = (int)((int64_t)a.n * (part + 1) / P);
    const int voff = chunk * C + vi;

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const int32_t* __restrict__ col = a.col;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    const XT* prev = static_cast<const XT*>(a.prev);
    XT* out = static_cast<XT*>(a.out);
    const XT* __restrict__ vin = static_cast<const XT*>(a.v);
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;
    const bool direct = vin != nullptr;

    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    const int stride = GW * RB;
    int r0 = lo + w * RB + wave * 4;      // first row of this wavefront's quad
    int row = r0 + slot;
    int64_t kb = 0;
    int cnt = 0;
    if (row < hi) {
        kb = rp[row];
        cnt = (int)(rp[row + 1] - kb);
    }
    int mc = 0;
    HT mh = zero_h<HT>();
    if (vi < cnt) {
        mc = col[kb + vi];
        mh = val[kb + vi];
    }
    for (; r0 < hi; r0 += stride) {
        const int nrow = r0 + stride + slot;
        int64_t nkb = 0;
        int ncnt = 0;
        if (nrow < hi) {
            nkb = rp[nrow];
            ncnt = (int)(rp[nrow + 1] - nkb);
        }
        XT own_c, own_p;
        zero(own_c);
        zero(own_p);
        if (row < hi) {
            const size_t idx = (size_t)row * R + voff;
            own_c = cur[idx];
            if (beta) own_p = prev[idx];
        }
        int mx = max(cnt, __shfl_xor(cnt, 16));
        mx = max(mx, __shfl_xor(mx, 32));
        mx = __builtin_amdgcn_readfirstlane(mx);
        AT acc;
        zero(acc);
        for (int k0 = 0; k0 < mx; k0 += 16) {
            // entries consumed after this chunk: the row's next 16, or the next quad's first 16
            int pc = 0;
            HT ph = zero_h<HT>();
            if (k0 + 16 < mx) {
                const int q = k0 + 16 + vi;
                if (q < cnt) {
                    pc = col[kb + q];
                    ph = val[kb + q];
                }
            } else if (vi < ncnt) {
                pc = col[nkb + vi];
                ph = val[nkb + vi];
            }
            const int m = cnt - k0;
            XT xs[16];
            xchunk_gather<XT, R, 0>(xs, m, mc, cur, voff);
            xchunk_mac<HT, XT, AT, 0>(acc, m, mh, xs);
            mc = pc;
            mh = ph;
        }
        if (mx == 0) {
            mc = 0;
            mh = zero_h<HT>();
            if (vi < ncnt) {
                mc = col[nkb + vi];
                mh = val[nkb + vi];
            }
        }
        if (row < hi) {
            const size_t idx = (size_t)row * R + voff;
            const XT c = own_c;
            const XT o = comb(alpha, acc, own_p);
            out[idx] = o;
            d1 += nrm2(c);
            d2 += redot(o, c);
            if (direct) d3 += redot(vin[idx], c);
        }
        row = nrow;
        kb = nkb;
        cnt = ncnt;
    }

    d1 += __shfl_xor(d1, 16);
    d2 += __shfl_xor(d2, 16);
    d3 += __shfl_xor(d3, 16);
    d1 += __shfl_xor(d1, 32);
    d2 += __shfl_xor(d2, 32);
    d3 += __shfl_xor(d3, 32);
    __shared__ double red[kWaves][C][3];
    if (slot == 0) {
        red[wave][vi][0] = d1;
        red[wave][vi][1] = d2;
        red[wave][vi][2] = d3;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        const int r = t / 3, k = t % 3;
        double sum = 0.0;
        if (r / C == chunk) {
            const int v = r % C;
            sum = red[0][v][k];
#pragma unroll
            for (int q = 1; q < kWaves; ++q) sum += red[q][v][k];
        }
        a.part[(size_t)lb * R * 3 + t] = sum;
    }
}

template <typename HT, typename XT>
hipError_t launch_xchunk(const StepArgs& a, hipStream_t s)
{
    switch (a.R) {
    case 16: hipLaunchKernelGGL((kpm_step_xchunk_kernel<HT, XT, 1>), dim3(a.grid), dim3(kThreads), 0, s, a); break;
    case 32: hipLaunchKernelGGL((kpm_step_xchunk_kernel<HT, XT, 2>), dim3(a.grid), dim3(kThreads), 0, s, a); break;
    case 64: hipLaunchKernelGGL((kpm_step_xchunk_kernel<HT, XT, 4>), dim3(a.grid), dim3(kThreads), 0, s, a); break;
    case 128: hipLaunchKernelGGL((kpm_step_xchunk_kernel<HT, XT, 8>), dim3(a.grid), dim3(kThreads), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// 16-byte lane elements: two complex FP32 values (float4) or one complex FP64 (double2)
__device__ inline void zero16(float4& a) { a = make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ inline void zero16(double2& a) { a = make_double2(0.0, 0.0); }
__device__ inline void split16(const float4& a, float2* x)
{
    x[0] = make_float2(a.x, a.y);
    x[1] = make_float2(a.z, a.w);
}
__device__ inline void split16(const double2& a, double2* x) { x[0] = a; }
__device__ inline float4 join16(const float2* x) { return make_float4(x[0].x, x[0].y, x[1].x, x[1].y); }
__device__ inline double2 join16(const double2* x) { return x[0]; }

// ---- variant 15: FP32 recursion with 16-byte gathers -----------------------------
// A complex FP32 row of T_j (R = 64) is 512 bytes.  The wavefront-per-row
// kernels gather it with 8-byte lanes, so a gather instruction moves half the
// bytes of the FP64 case for the same address work.  Here each lane carries a
// pair of vectors (16 bytes), 32 lanes cover a row and a wavefront owns two
// consecutive rows, one per half: one instruction gathers a neighbour row of
// each (1 KiB).  Both rows' entries come through the scalar path in
// BATCH-entry chunks (as in variant 8) and each half selects its own row's
// entry; every row is still summed entry by entry in CSR order.
template <typename HT, int BATCH>
__global__ __launch_bounds__(kThreads) void kpm_step_pair_kernel(const int64_t* __restrict__ rp,
                                                                 const int32_t* __restrict__ col,
                                                                 const HT* __restrict__ val,
                                                                 const float4* __restrict__ cur,
                                                                 float4* __restrict__ po,
                                                                 const float4* __restrict__ vin, int nrows,
                                                                 double alpha, int beta, double* __restrict__ part)
{
    constexpr int R = 64;
    constexpr int RQ = R / 2;             // float4 per row
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int half = lane >> 5;
    const int hl = lane & 31;             // vectors 2 hl and 2 hl + 1
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    const int stride = G * kWaves;
    const int npairs = (nrows + 1) / 2;
    const int64_t nnz = rp[nrows];        // chunk loads start at or before nnz (BATCH padding entries follow)
    const bool direct = vin != nullptr;
    const float al = (float)alpha;

    double d1a = 0.0, d1b = 0.0, d2a = 0.0, d2b = 0.0, d3a = 0.0, d3b = 0.0;
    // extents of pair pr: rows 2 pr and 2 pr + 1 (the second may not exist)
    auto extents = [&](int p, int64_t& qa, int64_t& ea, int64_t& qb, int64_t& eb) {
        qa = ea = qb = eb = 0;
        if (p < npairs) {
            const int ra = 2 * p;
            qa = rp[ra];
            ea = rp[ra + 1];
            qb = ea;
            eb = ra + 1 < nrows ? rp[ra + 2] : ea;
        }
    };
    int pr = __builtin_amdgcn_readfirstlane(lb * kWaves + wave);
    int64_t qA, eA, qB, eB;
    extents(pr, qA, eA, qB, eB);
    // whole-chunk loads: the arrays carry BATCH padding entries past nnz.  The
    // shorter row's chunk may run past its end (masked below); clamping keeps
    // every chunk start at or before nnz.  The column chunks that the next
    // gathers need (this pair's next chunk, or the next pair's first) are
    // loaded one chunk ahead, while the current chunk's gathers are in flight.
    ColChunk<BATCH> ca = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(qA, nnz));
    ColChunk<BATCH> cb = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(qB, nnz));
    for (; pr < npairs; pr += stride) {
        int64_t nqA, neA, nqB, neB;
        extents(pr + stride, nqA, neA, nqB, neB);
        const int rA = 2 * pr;
        const bool hasB = rA + 1 < nrows;
        const int row = half ? min(rA + 1, nrows - 1) : rA;
        const bool live = half == 0 || hasB;
        const size_t own = (size_t)row * RQ + hl;
        float4 oc = make_float4(0.f, 0.f, 0.f, 0.f), op = oc;
        if (live) {
            oc = cur[own];
            if (beta) op = po[own];
        }
        const int64_t my_end = half ? eB : eA;
        const int64_t my_beg = half ? qB : qA;
        const int64_t span = max(eA - qA, eB - qB);
        float2 a0 = make_float2(0.f, 0.f), a1 = a0;
        for (int64_t k = 0; k < span; k += BATCH) {
            const int64_t sA = min(qA + k, nnz), sB = min(qB + k, nnz);
            float4 x[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int c = half ? cb.c[u] : ca.c[u];
                x[u] = cur[(size_t)c * RQ + hl];
            }
            const ValChunk<HT, BATCH> ha = *reinterpret_cast<const ValChunk<HT, BATCH>*>(val + sA);
            const ValChunk<HT, BATCH> hb = *reinterpret_cast<const ValChunk<HT, BATCH>*>(val + sB);
            const bool more = k + BATCH < span;
            ca = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(more ? qA + k + BATCH : nqA, nnz));
            cb = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(more ? qB + k + BATCH : nqB, nnz));
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const HT h = half ? hb.h[u] : ha.h[u];
                if (my_beg + k + u < my_end) {
                    mac(a0, h, make_float2(x[u].x, x[u].y));
                    mac(a1, h, make_float2(x[u].z, x[u].w));
                }
            }
        }
        if (span == 0) {
            ca = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(nqA, nnz));
            cb = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(nqB, nnz));
        }
        if (live) {
            const float2 c0 = make_float2(oc.x, oc.y), c1 = make_float2(oc.z, oc.w);
            float2 o0, o1;
            o0.x = al * a0.x - op.x;
            o0.y = al * a0.y - op.y;
            o1.x = al * a1.x - op.z;
            o1.y = al * a1.y - op.w;
            po[own] = make_float4(o0.x, o0.y, o1.x, o1.y);
            d1a += nrm2(c0);
            d1b += nrm2(c1);
            d2a += redot(o0, c0);
            d2b += redot(o1, c1);
            if (direct) {
                const float4 vv = vin[own];
                d3a += redot(make_float2(vv.x, vv.y), c0);
                d3b += redot(make_float2(vv.z, vv.w), c1);
            }
        }
        qA = nqA;
        eA = neA;
        qB = nqB;
        eB = neB;
    }

    // fold the halves (fixed order), then the wavefronts
    d1a += __shfl_xor(d1a, 32);
    d1b += __shfl_xor(d1b, 32);
    d2a += __shfl_xor(d2a, 32);
    d2b += __shfl_xor(d2b, 32);
    d3a += __shfl_xor(d3a, 32);
    d3b += __shfl_xor(d3b, 32);
    __shared__ double red[kWaves][R][3];
    if (half == 0) {
        red[wave][2 * hl][0] = d1a;
        red[wave][2 * hl][1] = d2a;
        red[wave][2 * hl][2] = d3a;
        red[wave][2 * hl + 1][0] = d1b;
        red[wave][2 * hl + 1][1] = d2b;
        red[wave][2 * hl + 1][2] = d3b;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        const int r = t / 3, k = t % 3;
        double sum = red[0][r][k];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) sum += red[w][r][k];
        part[((size_t)lb * R + r) * 3 + k] = sum;
    }
}

template <typename HT, typename XT, int LPR, int VPL>
hipError_t launch_typed(const StepArgs& a, hipStream_t s)
{
    if (a.variant == 14) return launch_xchunk<HT, XT>(a, s);
    if constexpr (std::is_same<XT, float2>::value && sizeof(HT) <= 8) {
        if (a.variant == 15) {
            // complex FP32 vectors, R = 64, T_{j-1} and T_{j+1} in one buffer
            if (a.R != 64 || a.prev != a.out) return hipErrorInvalidValue;
            hipLaunchKernelGGL((kpm_step_pair_kernel<HT, 8>), dim3(a.grid), dim3(kThreads), 0, s, a.row_ptr, a.col,
                               static_cast<const HT*>(a.val), static_cast<const float4*>(a.cur),
                               static_cast<float4*>(a.out), static_cast<const float4*>(a.v), a.n, a.alpha, a.beta,
                               a.part);
            return hipGetLastError();
        }
    }
    if (LPR == 64 && a.variant == 1)
        hipLaunchKernelGGL((kpm_step_wave_kernel<HT, XT, VPL, 8>), dim3(a.grid), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL((kpm_step_kernel<HT, XT, LPR, VPL>), dim3(a.grid), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

template <typename HT, typename XT>
hipError_t launch_by_r(const StepArgs& a, hipStream_t s)
{
    switch (a.R) {
    case 1: return launch_typed<HT, XT, 1, 1>(a, s);
    case 2: return launch_typed<HT, XT, 2, 1>(a, s);
    case 4: return launch_typed<HT, XT, 4, 1>(a, s);
    case 8: return launch_typed<HT, XT, 8, 1>(a, s);
    case 16: return launch_typed<HT, XT, 16, 1>(a, s);
    case 32: return launch_typed<HT, XT, 32, 1>(a, s);
    case 64: return launch_typed<HT, XT, 64, 1>(a, s);
    case 128: return launch_typed<HT, XT, 64, 2>(a, s);
    default: return hipErrorInvalidValue;
    }
}

// ---- reduction of partials into moments ----------------------------------
__global__ __launch_bounds__(kThreads) void kpm_reduce_kernel(ReduceArgs a)
{
    const int r = blockIdx.x;
    const int R = a.R;
    __shared__ double s[3][kThreads];
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int g = threadIdx.x; g < a.grid; g += kThreads) {
        const double* p = a.part + ((size_t)g * R + r) * 3;
        t0 += p[0];
        t1 += p[1];
        t2 += p[2];
    }
    s[0][threadIdx.x] = t0;
    s[1][threadIdx.x] = t1;
    s[2][threadIdx.x] = t2;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
            s[2][threadIdx.x] += s[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double d1 = s[0][0], d2 = s[1][0], d3 = s[2][0];
    double* mu = a.mu + (size_t)r * a.M;
    const int j = a.j, M = a.M;
    auto load_re = [&](const void* buf, int row) -> double {
        const size_t idx = (size_t)row * R + r;
        switch (a.vec_type) {
        case VEC_REAL_F64: return static_cast<const double*>(buf)[idx];
        case VEC_CPX_F64: return static_cast<const double2*>(buf)[idx].x;
        case VEC_CPX_F32: return (double)static_cast<const float2*>(buf)[idx].x;
        default: return (double)static_cast<const float*>(buf)[idx];
        }
    };
    if (a.scheme == 1) {
        if (j == 0) { mu[0] = d1; mu[1] = d2; }
        else if (2 * j + 1 < M) { mu[2 * j] = 2.0 * d1 - mu[0]; mu[2 * j + 1] = 2.0 * d2 - mu[1]; }
        return;
    }
    if (j == 0) {
        if (a.scheme == 0) { mu[0] = 1.0; mu[1] = load_re(a.out, a.targets[r]); }
        else { mu[0] = d1; mu[1] = d2; }
        return;
    }
    if (j >= 2 && j < M / 2) {
        mu[j] = a.scheme == 0 ? load_re(a.cur, a.targets[r]) : d3;
        if (j >= M / 4) {
            mu[2 * j] = 2.0 * d1 - mu[0];
            mu[2 * j + 1] = 2.0 * d2 - mu[1];
        }
    }
}

// ---- start vectors ---------------------------------------------------------

__global__ void init_random_phase_kernel(void* x, int vec_type, int n, int R, uint64_t seed, int r0)
{
    const size_t total = (size_t)n * R;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const uint64_t i = t / R, r = t % R;
        const double phi = random_phase_angle(seed, r0 + r, i);
        double sn, cs;
        sincos(phi, &sn, &cs);
        if (vec_type == VEC_CPX_F64) static_cast<double2*>(x)[t] = make_double2(cs, sn);
        else static_cast<float2*>(x)[t] = make_float2((float)cs, (float)sn);
    }
}

__global__ void init_targets_kernel(void* x, int vec_type, int n, int R, const int32_t* targets)
{
    const size_t total = (size_t)n * R;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / R), r = (int)(t % R);
        const double v = (targets[r] == i) ? 1.0 : 0.0;
        switch (vec_type) {
        case VEC_REAL_F64: static_cast<double*>(x)[t] = v; break;
        case VEC_CPX_F64: static_cast<double2*>(x)[t] = make_double2(v, 0.0); break;
        case VEC_CPX_F32: static_cast<float2*>(x)[t] = make_float2((float)v, 0.0f); break;
        default: static_cast<float*>(x)[t] = (float)v; break;
        }
    }
}

// src: host-style [R][n] real or complex (interleaved) FP64, already on device
__global__ void init_from_host_layout_kernel(void* x, int vec_type, const double* src, int src_complex, int n, int R)
{
    const size_t total = (size_t)n * R;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t / R, r = t % R;
        double re_v, im_v;
        if (src_complex) {
            re_v = src[2 * (r * n + i)];
            im_v = src[2 * (r * n + i) + 1];
        } else {
            re_v = src[r * n + i];
            im_v = 0.0;
        }
        switch (vec_type) {
        case VEC_REAL_F64: static_cast<double*>(x)[t] = re_v; break;
        case VEC_CPX_F64: static_cast<double2*>(x)[t] = make_double2(re_v, im_v); break;
        case VEC_CPX_F32: static_cast<float2*>(x)[t] = make_float2((float)re_v, (float)im_v); break;
        default: static_cast<float*>(x)[t] = (float)re_v; break;
        }
    }
}

__global__ void convert_kernel(const double* in, float* out, int64_t count)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (float)in[t];
}

// ---- symmetric permutation H' = P H P^T (row q of H' = row iperm[q] of H) --
__global__ void perm_len_kernel(const int64_t* rp, const int32_t* iperm, int n, int64_t* len)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = iperm[q];
    len[q] = rp[i + 1] - rp[i];
}

template <typename V>
__global__ void perm_fill_kernel(const int64_t* rp, const int32_t* col, const V* val, const int32_t* iperm,
                                 const int32_t* perm, int n, const int64_t* prp, int32_t* pcol, V* pval)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = iperm[q];
    int64_t w = prp[q];
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e, ++w) {
        pcol[w] = perm[col[e]];
        pval[w] = val[e];
    }
}

__global__ void init_random_phase_perm_kernel(void* x, int vec_type, int n, int R, uint64_t seed, int r0,
                                              const int32_t* iperm)
{
    const size_t total = (size_t)n * R;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const uint64_t q = t / R, r = t % R;
        const uint64_t i = iperm ? (uint64_t)iperm[q] : q;
        const double phi = random_phase_angle(seed, r0 + r, i);
        double sn, cs;
        sincos(phi, &sn, &cs);
        if (vec_type == VEC_CPX_F64) static_cast<double2*>(x)[t] = make_double2(cs, sn);
        else static_cast<float2*>(x)[t] = make_float2((float)cs, (float)sn);
    }
}

__global__ void init_from_host_layout_perm_kernel(void* x, int vec_type, const double* src, int src_complex, int n,
                                                  int R, const int32_t* iperm)
{
    const size_t total = (size_t)n * R;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t q = t / R, r = t % R;
        const size_t i = iperm ? (size_t)iperm[q] : q;
        double re_v, im_v;
        if (src_complex) {
            re_v = src[2 * (r * n + i)];
            im_v = src[2 * (r * n + i) + 1];
        } else {
            re_v = src[r * n + i];
            im_v = 0.0;
        }
        switch (vec_type) {
        case VEC_REAL_F64: static_cast<double*>(x)[t] = re_v; break;
        case VEC_CPX_F64: static_cast<double2*>(x)[t] = make_double2(re_v, im_v); break;
        case VEC_CPX_F32: static_cast<float2*>(x)[t] = make_float2((float)re_v, (float)im_v); break;
        default: static_cast<float*>(x)[t] = (float)re_v; break;
        }
    }
}

inline dim3 stride_grid(size_t total)
{
    size_t g = (total + kThreads - 1) / kThreads;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

}  // namespace

hipError_t launch_kpm_step(int mat_type, int vec_type, const StepArgs& a, hipStream_t s)
{
    if (a.grid % 8 != 0 || a.grid <= 0) return hipErrorInvalidValue;
    if (mat_type == MAT_REAL_F64 && vec_type == VEC_REAL_F64) return launch_by_r<double, double>(a, s);
    if (mat_type == MAT_REAL_F64 && vec_type == VEC_CPX_F64) return launch_by_r<double, double2>(a, s);
    if (mat_type == MAT_CPX_F64 && vec_type == VEC_CPX_F64) return launch_by_r<double2, double2>(a, s);
    if (mat_type == MAT_REAL_F32 && vec_type == VEC_CPX_F32) return launch_by_r<float, float2>(a, s);
    if (mat_type == MAT_CPX_F32 && vec_type == VEC_CPX_F32) return launch_by_r<float2, float2>(a, s);
    if (mat_type == MAT_REAL_F32 && vec_type == VEC_REAL_F32) return launch_by_r<float, float>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_kpm_reduce(const ReduceArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(kpm_reduce_kernel, dim3(a.R), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_init_targets(void* x, int vec_type, int n, int R, const int32_t* targets, hipStream_t s)
{
    hipLaunchKernelGGL(init_targets_kernel, stride_grid((size_t)n * R), dim3(kThreads), 0, s, x, vec_type, n, R,
                       targets);
    return hipGetLastError();
}

hipError_t launch_init_random_phase(void* x, int vec_type, int n, int R, uint64_t seed, int r0, hipStream_t s)
{
    if (vec_type != VEC_CPX_F64 && vec_type != VEC_CPX_F32) return hipErrorInvalidValue;
    hipLaunchKernelGGL(init_random_phase_kernel, stride_grid((size_t)n * R), dim3(kThreads), 0, s, x, vec_type, n,
                       R, seed, r0);
    return hipGetLastError();
}

hipError_t launch_init_from_host_layout(void* x, int vec_type, const double* src, int src_complex, int n, int R,
                                        hipStream_t s)
{
    hipLaunchKernelGGL(init_from_host_layout_kernel, stride_grid((size_t)n * R), dim3(kThreads), 0, s, x, vec_type,
                       src, src_complex, n, R);
    return hipGetLastError();
}

hipError_t launch_init_random_phase_perm(void* x, int vec_type, int n, int R, uint64_t seed, int r0,
                                        const int32_t* iperm, hipStream_t s)
{
    if (vec_type != VEC_CPX_F64 && vec_type != VEC_CPX_F32) return hipErrorInvalidValue;
    hipLaunchKernelGGL(init_random_phase_perm_kernel, stride_grid((size_t)n * R), dim3(kThreads), 0, s, x, vec_type,
                       n, R, seed, r0, iperm);
    return hipGetLastError();
}

hipError_t launch_init_from_host_layout_perm(void* x, int vec_type, const double* src, int src_complex, int n, int R,
                                             const int32_t* iperm, hipStream_t s)
{
    hipLaunchKernelGGL(init_from_host_layout_perm_kernel, stride_grid((size_t)n * R), dim3(kThreads), 0, s, x,
                       vec_type, src, src_complex, n, R, iperm);
    return hipGetLastError();
}

hipError_t launch_perm_len(const int64_t* rp, const int32_t* iperm, int n, int64_t* len, hipStream_t s)
{
    hipLaunchKernelGGL(perm_len_kernel, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, rp, iperm, n, len);
    return hipGetLastError();
}

hipError_t launch_perm_fill(const int64_t* rp, const int32_t* col, const void* val, int val_bytes, const int32_t* iperm,
                            const int32_t* perm, int n, const int64_t* prp, int32_t* pcol, void* pval, hipStream_t s)
{
    const dim3 g((n + kThreads - 1) / kThreads);
    if (val_bytes == 16)
        hipLaunchKernelGGL(perm_fill_kernel<double2>, g, dim3(kThreads), 0, s, rp, col, (const double2*)val, iperm, perm,
                           n, prp, pcol, (double2*)pval);
    else if (val_bytes == 8)
        hipLaunchKernelGGL(perm_fill_kernel<double>, g, dim3(kThreads), 0, s, rp, col, (const double*)val, iperm, perm,
                           n, prp, pcol, (double*)pval);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_convert_f64_to_f32(const double* in, float* out, int64_t count, hipStream_t s)
{
    hipLaunchKernelGGL(convert_kernel, stride_grid((size_t)count), dim3(kThreads), 0, s, in, out, count);
    return hipGetLastError();
}

}  // namespace leo2d
