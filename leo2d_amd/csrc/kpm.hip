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

// sum over the 4 entry slots (lanes e, e+16, e+32, e+48)
__device__ inline void fold_slots(double& a)
{
    a += __shfl_xor(a, 16);
    a += __shfl_xor(a, 32);
}
__device__ inline void fold_slots(float& a)
{
    a += __shfl_xor(a, 16);
    a += __shfl_xor(a, 32);
}
__device__ inline void fold_slots(double2& a)
{
    fold_slots(a.x);
    fold_slots(a.y);
}
__device__ inline void fold_slots(float2& a)
{
    fold_slots(a.x);
    fold_slots(a.y);
}

// Streaming accesses (the matrix, T_{j-1} and T_{j+1}) can be issued with the
// non-temporal hint so that they do not displace the gathered T_j rows in L2.
typedef double dbl2v __attribute__((ext_vector_type(2)));
typedef float flt2v __attribute__((ext_vector_type(2)));
template <bool S> __device__ inline double ld_s(const double* p) { return S ? __builtin_nontemporal_load(p) : *p; }
template <bool S> __device__ inline float ld_s(const float* p) { return S ? __builtin_nontemporal_load(p) : *p; }
template <bool S> __device__ inline int ld_s(const int* p) { return S ? __builtin_nontemporal_load(p) : *p; }
template <bool S> __device__ inline double2 ld_s(const double2* p)
{
    if (!S) return *p;
    const dbl2v v = __builtin_nontemporal_load(reinterpret_cast<const dbl2v*>(p));
    return make_double2(v.x, v.y);
}
template <bool S> __device__ inline float2 ld_s(const float2* p)
{
    if (!S) return *p;
    const flt2v v = __builtin_nontemporal_load(reinterpret_cast<const flt2v*>(p));
    return make_float2(v.x, v.y);
}
template <bool S> __device__ inline void st_s(double* p, double v)
{
    if (S) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template <bool S> __device__ inline void st_s(float* p, float v)
{
    if (S) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template <bool S> __device__ inline void st_s(double2* p, double2 v)
{
    if (S) __builtin_nontemporal_store(dbl2v{v.x, v.y}, reinterpret_cast<dbl2v*>(p));
    else *p = v;
}
template <bool S> __device__ inline void st_s(float2* p, float2 v)
{
    if (S) __builtin_nontemporal_store(flt2v{v.x, v.y}, reinterpret_cast<flt2v*>(p));
    else *p = v;
}

template <typename HT, typename XT, int VPL, int BATCH, int NT = kThreads, bool STREAM = false>
__global__ __launch_bounds__(NT) void kpm_step_wave_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int NW = NT / 64;
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
    int my_c = (row < n && q0 + lane < q1) ? ld_s<STREAM>(col + q0 + lane) : 0;
    HT my_h = (row < n && q0 + lane < q1) ? ld_s<STREAM>(val + q0 + lane) : zero_h<HT>();

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
            c_reg = (base + lane < q1) ? ld_s<STREAM>(col + base + lane) : 0;
            h_reg = (base + lane < q1) ? ld_s<STREAM>(val + base + lane) : zero_h<HT>();
        }
        // next row's entries (issued before this row's epilogue traffic)
        nq0 = uniform64(nq0);
        nq1 = uniform64(nq1);
        if (next < n) {
            my_c = (nq0 + lane < nq1) ? ld_s<STREAM>(col + nq0 + lane) : 0;
            my_h = (nq0 + lane < nq1) ? ld_s<STREAM>(val + nq0 + lane) : zero_h<HT>();
        }
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            const size_t idx = (size_t)row * R + lane + 64 * v;
            const XT c = cur[idx];
            XT p;
            if (beta) p = ld_s<STREAM>(prev + idx);
            else zero(p);
            const XT o = comb(alpha, acc[v], p);
            st_s<STREAM>(out + idx, o);
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
    for (int t = threadIdx.x; t < R * 3; t += NT) {
        const int r = t / 3, k = t % 3;
        double s = red[0][r][k];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += red[w][r][k];
        a.part[((size_t)lb * R + r) * 3 + k] = s;
    }
}

// ---- LDS-tiled step (R a multiple of 16) --------------------------------------
// One workgroup owns one tile of consecutive (Morton-ordered, hence spatially
// compact) rows at a time.  For each group of 16 vectors it stages the tile's
// distinct neighbour rows x[U][16] into LDS with coalesced 256-byte (FP64)
// segments, then every row reads its nonzeros' x from LDS through the 16-bit
// tile-local column index.  L2 traffic per row drops from nnz_row x R x 16 B
// to (U / T) x R x 16 B (about 4-5 rows instead of 32 for bilayer graphene).
template <typename HT, typename XT, int NG>
__global__ __launch_bounds__(kThreads) void kpm_step_tiled_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int RG = kTileGroup;
    constexpr int RPW = 64 / RG;            // rows per wavefront pass
    constexpr int RPI = RPW * kWaves;       // rows per workgroup pass
    constexpr int SU = 8;                   // staging gathers in flight per thread
    constexpr int SLOTS = kThreads / RG;    // rows staged per workgroup pass
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int ucap = kTileLdsBytes / (RG * (int)sizeof(XT));
    XT* xs = reinterpret_cast<XT*>(lds_raw);
    int32_t* uc = reinterpret_cast<int32_t*>(lds_raw + kTileLdsBytes);   // tile column list
    const int R = RG * NG;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sub = lane / RG;
    const int e = lane % RG;
    const int slot = threadIdx.x / RG;      // staging row slot
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const uint16_t* __restrict__ lcol = a.lcol;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    const XT* prev = static_cast<const XT*>(a.prev);
    XT* out = static_cast<XT*>(a.out);
    const XT* __restrict__ vin = static_cast<const XT*>(a.v);
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;
    const bool direct = vin != nullptr;

    double d1[NG], d2[NG], d3[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) { d1[g] = 0.0; d2[g] = 0.0; d3[g] = 0.0; }

    for (int tile = lb; tile < a.ntiles; tile += G) {
        const int r_begin = a.tile_row0[tile], r_end = a.tile_row0[tile + 1];
        const int64_t u_begin = a.tile_u0[tile];
        const int U = min((int)(a.tile_u0[tile + 1] - u_begin), ucap);
        __syncthreads();
        for (int u = threadIdx.x; u < U; u += kThreads) uc[u] = a.ucols[u_begin + u];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            __syncthreads();
            for (int m0 = slot; m0 < U; m0 += SLOTS * SU) {
                XT xv[SU];
#pragma unroll
                for (int m = 0; m < SU; ++m) {
                    const int u = m0 + m * SLOTS;
                    if (u < U) xv[m] = cur[(size_t)uc[u] * R + g * RG + e];
                }
#pragma unroll
                for (int m = 0; m < SU; ++m) {
                    const int u = m0 + m * SLOTS;
                    if (u < U) xs[u * RG + e] = xv[m];
                }
            }
            __syncthreads();
            for (int row0 = r_begin + wave * RPW; row0 < r_end; row0 += RPI) {
                const int row = row0 + sub;
                if (row >= r_end) continue;
                AT acc;
                zero(acc);
                int64_t q = rp[row];
                const int64_t qe = rp[row + 1];
                for (; q + 4 <= qe; q += 4) {
                    const HT h0 = val[q], h1 = val[q + 1], h2 = val[q + 2], h3 = val[q + 3];
                    const int l0 = lcol[q], l1 = lcol[q + 1], l2 = lcol[q + 2], l3 = lcol[q + 3];
                    const XT x0 = xs[l0 * RG + e];
                    const XT x1 = xs[l1 * RG + e];
                    const XT x2 = xs[l2 * RG + e];
                    const XT x3 = xs[l3 * RG + e];
                    mac(acc, h0, x0);
                    mac(acc, h1, x1);
                    mac(acc, h2, x2);
                    mac(acc, h3, x3);
                }
                for (; q < qe; ++q) mac(acc, val[q], xs[lcol[q] * RG + e]);
                const size_t idx = (size_t)row * R + g * RG + e;
                const XT c = cur[idx];
                XT p;
                if (beta) p = prev[idx];
                else zero(p);
                const XT o = comb(alpha, acc, p);
                out[idx] = o;
                d1[g] += nrm2(c);
                d2[g] += redot(o, c);
                if (direct) d3[g] += redot(vin[idx], c);
            }
        }
    }

    // fold the row slots of a wavefront, then the wavefronts (LDS reused)
#pragma unroll
    for (int off = RG; off < 64; off <<= 1) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            d1[g] += __shfl_xor(d1[g], off);
            d2[g] += __shfl_xor(d2[g], off);
            d3[g] += __shfl_xor(d3[g], off);
        }
    }
    __syncthreads();
    double* red = reinterpret_cast<double*>(lds_raw);   // [kWaves][R][3]
    if (sub == 0) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int r = g * RG + e;
            red[(wave * R + r) * 3 + 0] = d1[g];
            red[(wave * R + r) * 3 + 1] = d2[g];
            red[(wave * R + r) * 3 + 2] = d3[g];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        double s = red[t];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += red[w * R * 3 + t];
        a.part[(size_t)lb * R * 3 + t] = s;
    }
}

// Variant 4: same tiles and LDS staging, but one row per wavefront pass: lane
// = (entry slot k in 0..3) x (vector e in 0..15).  Lane group k owns entries
// k, k+4, ..., k+28 of the row, loaded together up front (and prefetched for
// the wavefront's next row), so a row costs one memory round trip; the four
// partial sums are folded with two cross-lane shuffles.
template <typename HT, typename XT, int NG>
__global__ __launch_bounds__(kThreads) void kpm_step_tiled_row_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int RG = kTileGroup;
    constexpr int KS = 64 / RG;             // entry slots per wavefront (4)
    constexpr int EPL = 8;                  // entries per lane per chunk
    constexpr int CH = KS * EPL;            // entries per chunk (32)
    constexpr int SU = 8;
    constexpr int SLOTS = kThreads / RG;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int ucap = kTileLdsBytes / (RG * (int)sizeof(XT));
    XT* xs = reinterpret_cast<XT*>(lds_raw);
    int32_t* uc = reinterpret_cast<int32_t*>(lds_raw + kTileLdsBytes);
    const int R = RG * NG;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int k = lane / RG;
    const int e = lane % RG;
    const int slot = threadIdx.x / RG;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const uint16_t* __restrict__ lcol = a.lcol;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    const XT* prev = static_cast<const XT*>(a.prev);
    XT* out = static_cast<XT*>(a.out);
    const XT* __restrict__ vin = static_cast<const XT*>(a.v);
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;
    const bool direct = vin != nullptr;

    double d1[NG], d2[NG], d3[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) { d1[g] = 0.0; d2[g] = 0.0; d3[g] = 0.0; }

    for (int tile = lb; tile < a.ntiles; tile += G) {
        const int r_begin = a.tile_row0[tile], r_end = a.tile_row0[tile + 1];
        const int64_t u_begin = a.tile_u0[tile];
        const int U = min((int)(a.tile_u0[tile + 1] - u_begin), ucap);
        __syncthreads();
        for (int u = threadIdx.x; u < U; u += kThreads) uc[u] = a.ucols[u_begin + u];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            __syncthreads();
            for (int m0 = slot; m0 < U; m0 += SLOTS * SU) {
                XT xv[SU];
#pragma unroll
                for (int m = 0; m < SU; ++m) {
                    const int u = m0 + m * SLOTS;
                    if (u < U) xv[m] = cur[(size_t)uc[u] * R + g * RG + e];
                }
#pragma unroll
                for (int m = 0; m < SU; ++m) {
                    const int u = m0 + m * SLOTS;
                    if (u < U) xs[u * RG + e] = xv[m];
                }
            }
            __syncthreads();
            int row = r_begin + wave;
            HT hv[EPL];
            int lv[EPL];
            int64_t q0 = 0, q1 = 0;
            if (row < r_end) {
                q0 = rp[row];
                q1 = rp[row + 1];
#pragma unroll
                for (int i = 0; i < EPL; ++i) {
                    const int64_t q = q0 + k + KS * i;
                    hv[i] = q < q1 ? val[q] : zero_h<HT>();
                    lv[i] = q < q1 ? (int)lcol[q] : 0;
                }
            }
            for (; row < r_end; row += kWaves) {
                // prefetch the next row's first chunk
                const int nrow = row + kWaves;
                HT hn[EPL];
                int ln[EPL];
                int64_t n0 = 0, n1 = 0;
                if (nrow < r_end) {
                    n0 = rp[nrow];
                    n1 = rp[nrow + 1];
#pragma unroll
                    for (int i = 0; i < EPL; ++i) {
                        const int64_t q = n0 + k + KS * i;
                        hn[i] = q < n1 ? val[q] : zero_h<HT>();
                        ln[i] = q < n1 ? (int)lcol[q] : 0;
                    }
                }
                AT acc;
                zero(acc);
                const int64_t cnt = q1 - q0;
#pragma unroll
                for (int i = 0; i < EPL; ++i)
                    if (k + KS * i < cnt) mac(acc, hv[i], xs[lv[i] * RG + e]);
                for (int64_t base = q0 + CH; base < q1; base += CH) {   // rows longer than one chunk
#pragma unroll
                    for (int i = 0; i < EPL; ++i) {
                        const int64_t q = base + k + KS * i;
                        if (q < q1) mac(acc, val[q], xs[(int)lcol[q] * RG + e]);
                    }
                }
                // fold the four entry slots
                fold_slots(acc);
                if (k == 0) {
                    const size_t idx = (size_t)row * R + g * RG + e;
                    const XT c = cur[idx];
                    XT p;
                    if (beta) p = prev[idx];
                    else zero(p);
                    const XT o = comb(alpha, acc, p);
                    out[idx] = o;
                    d1[g] += nrm2(c);
                    d2[g] += redot(o, c);
                    if (direct) d3[g] += redot(vin[idx], c);
                }
#pragma unroll
                for (int i = 0; i < EPL; ++i) { hv[i] = hn[i]; lv[i] = ln[i]; }
                q0 = n0;
                q1 = n1;
            }
        }
    }

    __syncthreads();
    double* red = reinterpret_cast<double*>(lds_raw);   // [kWaves][R][3]
    if (k == 0) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int r = g * RG + e;
            red[(wave * R + r) * 3 + 0] = d1[g];
            red[(wave * R + r) * 3 + 1] = d2[g];
            red[(wave * R + r) * 3 + 2] = d3[g];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        double s = red[t];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += red[w * R * 3 + t];
        a.part[(size_t)lb * R * 3 + t] = s;
    }
}

template <typename HT, typename XT, int NG>
hipError_t launch_tiled_row(const StepArgs& a, hipStream_t s)
{
    const size_t shmem = (size_t)kTileLdsBytes + sizeof(int32_t) * (kTileLdsBytes / (kTileGroup * sizeof(XT)));
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)kpm_step_tiled_row_kernel<HT, XT, NG>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((kpm_step_tiled_row_kernel<HT, XT, NG>), dim3(a.grid), dim3(kThreads), shmem, s, a);
    return hipGetLastError();
}

// Variant 5: tiles staged with all R vectors (R = 64 * VPL) and one row per
// wavefront (lane = vector).  The row's (tile-local column, value) entries are
// fetched with one coalesced load per 64 entries and broadcast with readlane,
// so the values feed v_fma as scalar operands and x comes from LDS: the vector
// memory pipe only carries the staging loads (about U/T rows per row) and the
// row's own T_j, T_{j-1} and T_{j+1}.
template <typename HT, typename XT, int VPL, int SU>
__global__ __launch_bounds__(kThreads) void kpm_step_tiled_wave_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int R = 64 * VPL;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int ucap = a.tile_lds_bytes / (R * (int)sizeof(XT));
    XT* xs = reinterpret_cast<XT*>(lds_raw);
    int32_t* uc = reinterpret_cast<int32_t*>(lds_raw + a.tile_lds_bytes);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const uint16_t* __restrict__ lcol = a.lcol;
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

    for (int tile = lb; tile < a.ntiles; tile += G) {
        const int r_begin = a.tile_row0[tile], r_end = a.tile_row0[tile + 1];
        const int64_t u_begin = a.tile_u0[tile];
        const int U = min((int)(a.tile_u0[tile + 1] - u_begin), ucap);
        __syncthreads();
        for (int u = threadIdx.x; u < U; u += kThreads) uc[u] = a.ucols[u_begin + u];
        __syncthreads();
        // stage x[U][R]: wavefront w takes staged rows w, w+4, ...
        for (int u0 = wave; u0 < U; u0 += kWaves * SU) {
            XT xv[SU][VPL];
#pragma unroll
            for (int m = 0; m < SU; ++m) {
                const int u = u0 + m * kWaves;
                if (u < U) {
                    const size_t base = (size_t)uc[u] * R + lane;
#pragma unroll
                    for (int v = 0; v < VPL; ++v) xv[m][v] = cur[base + 64 * v];
                }
            }
#pragma unroll
            for (int m = 0; m < SU; ++m) {
                const int u = u0 + m * kWaves;
                if (u < U) {
#pragma unroll
                    for (int v = 0; v < VPL; ++v) xs[u * R + lane + 64 * v] = xv[m][v];
                }
            }
        }
        __syncthreads();
        int row = r_begin + wave;
        int64_t q0 = 0, q1 = 0;
        int l_reg = 0;
        HT h_reg = zero_h<HT>();
        if (row < r_end) {
            q0 = uniform64(rp[row]);
            q1 = uniform64(rp[row + 1]);
            if (q0 + lane < q1) {
                l_reg = lcol[q0 + lane];
                h_reg = val[q0 + lane];
            }
        }
        for (; row < r_end; row += kWaves) {
            const int nrow = row + kWaves;
            int64_t n0 = 0, n1 = 0;
            if (nrow < r_end) {
                n0 = rp[nrow];
                n1 = rp[nrow + 1];
            }
            AT acc[VPL];
#pragma unroll
            for (int v = 0; v < VPL; ++v) zero(acc[v]);
            int64_t base = q0;
            for (;;) {
                const int cnt = (int)min((int64_t)64, q1 - base);
                int kk = 0;
                for (; kk + 4 <= cnt; kk += 4) {
                    const int c0 = rl(l_reg, kk), c1 = rl(l_reg, kk + 1), c2 = rl(l_reg, kk + 2), c3 = rl(l_reg, kk + 3);
                    const HT h0 = rl(h_reg, kk), h1 = rl(h_reg, kk + 1), h2 = rl(h_reg, kk + 2), h3 = rl(h_reg, kk + 3);
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        const XT x0 = xs[c0 * R + lane + 64 * v];
                        const XT x1 = xs[c1 * R + lane + 64 * v];
                        const XT x2 = xs[c2 * R + lane + 64 * v];
                        const XT x3 = xs[c3 * R + lane + 64 * v];
                        mac(acc[v], h0, x0);
                        mac(acc[v], h1, x1);
                        mac(acc[v], h2, x2);
                        mac(acc[v], h3, x3);
                    }
                }
                for (; kk < cnt; ++kk) {
                    const int c0 = rl(l_reg, kk);
                    const HT h0 = rl(h_reg, kk);
#pragma unroll
                    for (int v = 0; v < VPL; ++v) mac(acc[v], h0, xs[c0 * R + lane + 64 * v]);
                }
                base += 64;
                if (base >= q1) break;
                l_reg = (base + lane < q1) ? (int)lcol[base + lane] : 0;
                h_reg = (base + lane < q1) ? val[base + lane] : zero_h<HT>();
            }
            n0 = uniform64(n0);
            n1 = uniform64(n1);
            if (nrow < r_end) {
                l_reg = (n0 + lane < n1) ? (int)lcol[n0 + lane] : 0;
                h_reg = (n0 + lane < n1) ? val[n0 + lane] : zero_h<HT>();
            }
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                const size_t idx = (size_t)row * R + lane + 64 * v;
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
            q0 = n0;
            q1 = n1;
        }
    }

    __syncthreads();
    double* red = reinterpret_cast<double*>(lds_raw);   // [kWaves][R][3]
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        const int r = lane + 64 * v;
        red[(wave * R + r) * 3 + 0] = d1[v];
        red[(wave * R + r) * 3 + 1] = d2[v];
        red[(wave * R + r) * 3 + 2] = d3[v];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        double s = red[t];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += red[w * R * 3 + t];
        a.part[(size_t)lb * R * 3 + t] = s;
    }
}

template <typename HT, typename XT, int VPL, int SU>
hipError_t launch_tiled_wave_su(const StepArgs& a, hipStream_t s)
{
    const size_t shmem = (size_t)a.tile_lds_bytes + sizeof(int32_t) * (a.tile_lds_bytes / (64 * VPL * sizeof(XT)));
    static int attr_bytes = 0;
    if (attr_bytes < (int)shmem) {
        hipError_t e = hipFuncSetAttribute((const void*)kpm_step_tiled_wave_kernel<HT, XT, VPL, SU>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
        if (e != hipSuccess) return e;
        attr_bytes = (int)shmem;
    }
    hipLaunchKernelGGL((kpm_step_tiled_wave_kernel<HT, XT, VPL, SU>), dim3(a.grid), dim3(kThreads), shmem, s, a);
    return hipGetLastError();
}

// all of a tile's staged rows are loaded in one burst: SU >= ucap / 4 rows per wavefront
template <typename HT, typename XT, int VPL>
hipError_t launch_tiled_wave(const StepArgs& a, hipStream_t s)
{
    const int ucap = a.tile_lds_bytes / (64 * VPL * (int)sizeof(XT));
    const int per_wave = (ucap + kWaves - 1) / kWaves;
    if (per_wave <= 8) return launch_tiled_wave_su<HT, XT, VPL, 8>(a, s);
    if (per_wave <= 18) return launch_tiled_wave_su<HT, XT, VPL, 18>(a, s);
    if (per_wave <= 36) return launch_tiled_wave_su<HT, XT, VPL, 36>(a, s);
    return launch_tiled_wave_su<HT, XT, VPL, 72>(a, s);
}

template <int B> struct alignas(4) ColChunk { int c[B]; };
template <typename HT, int B> struct ValChunk { HT h[B]; };

// Variant 8: wavefront per row, H through the scalar path.  The row index is
// wave-uniform and every pointer is restrict-qualified (T_{j-1} and T_{j+1}
// share one pointer), so the compiler fetches the row's column indices and
// values with scalar loads into SGPRs: no VALU broadcasts, no vector-memory
// traffic for H; the vector pipe carries only the x gathers (SGPR row base +
// lane offset) and the row's own T_j / T_{j-1} / T_{j+1}.
template <typename HT, typename XT, int VPL, int BATCH>
__global__ __launch_bounds__(kThreads) void kpm_step_smem_kernel(const int64_t* __restrict__ rp,
                                                                 const int32_t* __restrict__ col,
                                                                 const HT* __restrict__ val,
                                                                 const XT* __restrict__ cur,
                                                                 XT* __restrict__ po,
                                                                 const XT* __restrict__ vin, int nrows, double alpha,
                                                                 int beta, double* __restrict__ part, int xcd_rows)
{
    using AT = typename acc_of<XT>::type;
    constexpr int R = 64 * VPL;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    int first, n, stride;
    if (xcd_rows > 0) {
        first = (b & 7) * xcd_rows + (b >> 3) * kWaves;
        n = min(nrows, ((b & 7) + 1) * xcd_rows);
        stride = (G >> 3) * kWaves;
    } else {
        first = lb * kWaves;
        n = nrows;
        stride = G * kWaves;
    }
    const bool direct = vin != nullptr;

    double d1[VPL], d2[VPL], d3[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) { d1[v] = 0.0; d2[v] = 0.0; d3[v] = 0.0; }

    // whole-chunk loads: the arrays carry BATCH padding entries (column 0,
    // value 0) past nnz and no chunk starts beyond nnz.  The column chunk that
    // the next gathers need (the row's next chunk, or the next row's first) is
    // loaded one chunk ahead, while the current chunk's gathers are in flight.
    const int64_t nnz = rp[nrows];
    int row = __builtin_amdgcn_readfirstlane(first + wave);
    int64_t q0 = 0, q1 = 0;
    if (row < n) {
        q0 = rp[row];
        q1 = rp[row + 1];
    }
    ColChunk<BATCH> cc = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(q0, nnz));
    for (; row < n; row += stride) {
        const int next = row + stride;
        int64_t nq0 = 0, nq1 = 0;
        if (next < n) {
            nq0 = rp[next];
            nq1 = rp[next + 1];
        }
        // the row's own T_j and T_{j-1} elements, issued ahead of the gathers
        XT own_c[VPL], own_p[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            const size_t idx = (size_t)row * R + lane + 64 * v;
            own_c[v] = cur[idx];
            if (beta) own_p[v] = po[idx];
            else zero(own_p[v]);
        }
        AT acc[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) zero(acc[v]);
        for (int64_t q = q0; q < q1; q += BATCH) {
            XT x[BATCH][VPL];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const XT* xr = cur + (size_t)cc.c[u] * R;
#pragma unroll
                for (int v = 0; v < VPL; ++v) x[u][v] = xr[lane + 64 * v];
            }
            const ValChunk<HT, BATCH> hc = *reinterpret_cast<const ValChunk<HT, BATCH>*>(val + q);
            const int64_t nxt = q + BATCH < q1 ? q + BATCH : nq0;
            cc = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(nxt, nnz));
#pragma unroll
            for (int u = 0; u < BATCH; ++u)
                if (q + u < q1) {
#pragma unroll
                    for (int v = 0; v < VPL; ++v) mac(acc[v], hc.h[u], x[u][v]);
                }
        }
        if (q0 == q1) cc = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(nq0, nnz));
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            const size_t idx = (size_t)row * R + lane + 64 * v;
            const XT cv = own_c[v];
            const XT o = comb(alpha, acc[v], own_p[v]);
            po[idx] = o;
            d1[v] += nrm2(cv);
            d2[v] += redot(o, cv);
            if (direct) d3[v] += redot(vin[idx], cv);
        }
        q0 = nq0;
        q1 = nq1;
    }

    __shared__ double red[kWaves][R][3];
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
        for (int w = 1; w < kWaves; ++w) s += red[w][r][k];
        part[((size_t)lb * R + r) * 3 + k] = s;
    }
}

// ---- variant 11: LDS-tiled step, entries broadcast within 16-lane rows ------------
// Rows are grouped into tiles (at most kTileMaxRows consecutive rows whose
// distinct columns fit the LDS budget).  For each group of 16 vectors, the
// workgroup stages the tile's distinct neighbour rows (16 x element each, one
// 256-byte run for complex FP64) in LDS once, then every row reads its
// neighbours from LDS: the L2 sees each gathered row once per tile instead of
// once per referencing row.  Lane = (row slot 0..3) x (vector 0..15): the 16
// lanes of a row load 16 of its (local column, value) entries in one
// coalesced access and pass them to each other with DPP row broadcasts
// (row_newbcast), so no entry load sits on the dependent path of an LDS read.
// The first chunk of a slot's next row and its T_j / T_{j-1} elements are
// loaded before the current row is computed.  Several workgroups share a CU
// (LDS budget set at launch), so staging in one overlaps compute in another.
// acc += sum over the chunk's first `cnt` entries (entry kk held by lane kk of the 16-lane row)
template <typename HT, typename XT, typename AT, int K>
__device__ inline void chunk_mac(AT& acc, int cnt, int mc, HT mh, const XT* xs, int e)
{
    if constexpr (K < 16) {
        if (K < cnt) {
            const int c = bc16<K>(mc);
            const HT h = bc16<K>(mh);
            mac(acc, h, xs[c * 16 + e]);
            chunk_mac<HT, XT, AT, K + 1>(acc, cnt, mc, mh, xs, e);
        }
    }
}

template <typename HT, typename XT, int NG>
__global__ __launch_bounds__(kThreads) void kpm_step_tiled_dpp_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int RG = 16;
    constexpr int SU = 8;                    // staging loads in flight per thread
    constexpr int SLOTS = kThreads / RG;     // rows staged per pass
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int ucap = a.tile_lds_bytes / (RG * (int)sizeof(XT));
    XT* xs = reinterpret_cast<XT*>(lds_raw);
    const int R = RG * NG;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sub = lane >> 4;               // row slot in the wavefront
    const int e = lane & 15;                 // vector in the group
    const int slot = threadIdx.x >> 4;       // staging slot
    const int rslot = wave * 4 + sub;        // row slot in the workgroup (0..15)
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const uint16_t* __restrict__ lcol = a.lcol;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    const XT* prev = static_cast<const XT*>(a.prev);
    XT* out = static_cast<XT*>(a.out);
    const XT* __restrict__ vin = static_cast<const XT*>(a.v);
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;
    const bool direct = vin != nullptr;

    double d1[NG], d2[NG], d3[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) { d1[g] = 0.0; d2[g] = 0.0; d3[g] = 0.0; }

    for (int tile = lb; tile < a.ntiles; tile += G) {
        const int r_begin = a.tile_row0[tile], r_end = a.tile_row0[tile + 1];
        const int64_t u_begin = a.tile_u0[tile];
        const int U = min((int)(a.tile_u0[tile + 1] - u_begin), ucap);
#pragma unroll 1
        for (int g = 0; g < NG; ++g) {
            __syncthreads();   // every row of the previous group has read xs
            for (int m0 = slot; m0 < U; m0 += SLOTS * SU) {
                XT xv[SU];
#pragma unroll
                for (int m = 0; m < SU; ++m) {
                    const int u = m0 + m * SLOTS;
                    if (u < U) xv[m] = cur[(size_t)a.ucols[u_begin + u] * R + g * RG + e];
                }
#pragma unroll
                for (int m = 0; m < SU; ++m) {
                    const int u = m0 + m * SLOTS;
                    if (u < U) xs[u * RG + e] = xv[m];
                }
            }
            __syncthreads();
            // rows r_begin + rslot, + 16, ...; first chunk of the first row
            int row = r_begin + rslot;
            int64_t q0 = 0, q1 = 0;
            int mc = 0;
            HT mh = zero_h<HT>();
            if (row < r_end) {
                q0 = rp[row];
                q1 = rp[row + 1];
                if (q0 + e < q1) {
                    mc = lcol[q0 + e];
                    mh = val[q0 + e];
                }
            }
            for (; row < r_end; row += 16) {
                const size_t idx = (size_t)row * R + g * RG + e;
                const XT cv = cur[idx];
                XT pv;
                if (beta) pv = prev[idx];
                else zero(pv);
                const XT vv = direct ? vin[idx] : cv;
                // next row of this slot: first chunk in flight during this row
                const int nrow = row + 16;
                int64_t n0 = 0, n1 = 0;
                int nmc = 0;
                HT nmh = zero_h<HT>();
                if (nrow < r_end) {
                    n0 = rp[nrow];
                    n1 = rp[nrow + 1];
                    if (n0 + e < n1) {
                        nmc = lcol[n0 + e];
                        nmh = val[n0 + e];
                    }
                }
                AT acc;
                zero(acc);
                for (int64_t base = q0;;) {
                    const int cnt = (int)min((int64_t)16, q1 - base);
                    chunk_mac<HT, XT, AT, 0>(acc, cnt, mc, mh, xs, e);
                    base += 16;
                    if (base >= q1) break;
                    mc = (base + e < q1) ? (int)lcol[base + e] : 0;
                    mh = (base + e < q1) ? val[base + e] : zero_h<HT>();
                }
                const XT o = comb(alpha, acc, pv);
                out[idx] = o;
                d1[g] += nrm2(cv);
                d2[g] += redot(o, cv);
                if (direct) d3[g] += redot(vv, cv);
                q0 = n0;
                q1 = n1;
                mc = nmc;
                mh = nmh;
            }
        }
    }

    // fold the row slots of a wavefront, then the wavefronts (LDS reused)
#pragma unroll
    for (int off = RG; off < 64; off <<= 1) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            d1[g] += __shfl_xor(d1[g], off);
            d2[g] += __shfl_xor(d2[g], off);
            d3[g] += __shfl_xor(d3[g], off);
        }
    }
    __syncthreads();
    double* red = reinterpret_cast<double*>(lds_raw);   // [kWaves][R][3]
    if (sub == 0) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int r = g * RG + e;
            red[(wave * R + r) * 3 + 0] = d1[g];
            red[(wave * R + r) * 3 + 1] = d2[g];
            red[(wave * R + r) * 3 + 2] = d3[g];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += kThreads) {
        double sum = red[t];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) sum += red[w * R * 3 + t];
        a.part[(size_t)lb * R * 3 + t] = sum;
    }
}

template <typename HT, typename XT, int NG>
hipError_t launch_tiled_dpp(const StepArgs& a, hipStream_t s)
{
    // LDS: staged rows, at least the [kWaves][R][3] doubles of the final fold
    const size_t shmem = std::max((size_t)a.tile_lds_bytes, sizeof(double) * kWaves * 16 * NG * 3);
    hipError_t e = hipFuncSetAttribute((const void*)kpm_step_tiled_dpp_kernel<HT, XT, NG>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((kpm_step_tiled_dpp_kernel<HT, XT, NG>), dim3(a.grid), dim3(kThreads), shmem, s, a);
    return hipGetLastError();
}

// ---- variant 13: vector chunks over LDS-staged row tiles ----------------------
// The step's bound is the gather stream: every T_j row is fetched once per
// nonzero that points at it (about 32 times per step on bilayer graphene).
// This kernel arranges those fetches so that most of them hit in the CU's L1:
//  * a workgroup owns a contiguous run of row tiles (in the spatial order a
//    compact patch of the lattice) and walks it tile by tile;
//  * a tile's CSR entries (and tile-relative row offsets) are staged in LDS once;
//  * the tile is then swept once per chunk of C vectors, C chosen so that one
//    row's chunk is a single 128-byte line (C = 8 for complex FP64, 16 for
//    complex FP32).  A lane owns one (row, vector) pair and a wavefront covers
//    64 / C rows, so a gather instruction fetches 64 / C whole lines.
// Within one sweep the tile's distinct neighbour lines (about 3 per row) stay
// in L1, so the L2 request stream falls from ~32 lines per row and chunk to ~3.
// Partials: each sweep folds its per-lane sums over the wavefront's row slots
// (fixed xor tree) into this wavefront's LDS slot; the end folds the
// wavefronts in order.  The tile order, sweep order and trees are fixed, so the
// moments are deterministic.
template <typename HT, typename XT, int C, int NT, int NCH>
__global__ __launch_bounds__(NT) void kpm_step_chunk_kernel(StepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int RPW = 64 / C;          // rows per wavefront
    constexpr int NW = NT / 64;
    constexpr int R = C * NCH;
    constexpr int U = 8;                 // gathers in flight per lane
    extern __shared__ __align__(16) unsigned char lds[];
    double* red = reinterpret_cast<double*>(lds);                                   // [NW][R][3]
    HT* s_val = reinterpret_cast<HT*>(lds + sizeof(double) * NW * R * 3);           // [ecap]
    int32_t* s_col = reinterpret_cast<int32_t*>(s_val + a.chunk_ecap);              // [ecap]
    int32_t* s_rp = s_col + a.chunk_ecap;                                           // [tmax + 1]

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int rsub = lane / C;
    const int vi = lane % C;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    const int t_begin = a.chunk_wg[lb];
    const int t_end = a.chunk_wg[lb + 1];

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

    for (int t = lane; t < R * 3; t += 64) red[wave * R * 3 + t] = 0.0;

    for (int t = t_begin; t < t_end; ++t) {
        const int r0 = a.chunk_row0[t];
        const int nr = a.chunk_row0[t + 1] - r0;
        const int64_t e0 = rp[r0];
        const int ne = (int)(rp[r0 + nr] - e0);
        __syncthreads();   // the previous tile's entries are no longer read
        for (int q = threadIdx.x; q < ne; q += NT) {
            s_col[q] = col[e0 + q];
            s_val[q] = val[e0 + q];
        }
        for (int q = threadIdx.x; q <= nr; q += NT) s_rp[q] = (int)(rp[r0 + q] - e0);
        __syncthreads();
        for (int ch = 0; ch < NCH; ++ch) {
            const int voff = ch * C + vi;
            double d1 = 0.0, d2 = 0.0, d3 = 0.0;
            for (int rb = 0; rb < nr; rb += NW * RPW) {
                const int lr = rb + wave * RPW + rsub;
                if (lr < nr) {
                    const int kb = s_rp[lr], ke = s_rp[lr + 1];
                    AT acc;
                    zero(acc);
                    for (int k = kb; k < ke; k += U) {
                        XT x[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int kk = min(k + u, ke - 1);
                            x[u] = cur[(size_t)s_col[kk] * R + voff];
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (k + u < ke) mac(acc, s_val[k + u], x[u]);
                    }
                    const size_t idx = (size_t)(r0 + lr) * R + voff;
                    const XT c = cur[idx];
                    XT p;
                    if (beta) p = prev[idx];
                    else zero(p);
                    const XT o = comb(alpha, acc, p);
                    out[idx] = o;
                    d1 += nrm2(c);
                    d2 += redot(o, c);
                    if (direct) d3 += redot(vin[idx], c);
                }
            }
#pragma unroll
            for (int m = C; m < 64; m <<= 1) {
                d1 += __shfl_xor(d1, m);
                d2 += __shfl_xor(d2, m);
                d3 += __shfl_xor(d3, m);
            }
            if (rsub == 0) {
                double* slot = red + ((size_t)wave * R + ch * C + vi) * 3;
                slot[0] += d1;
                slot[1] += d2;
                slot[2] += d3;
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * 3; t += NT) {
        double sum = red[t];
#pragma unroll
        for (int w = 1; w < NW; ++w) sum += red[w * R * 3 + t];
        a.part[(size_t)lb * R * 3 + t] = sum;
    }
}

template <typename HT, typename XT, int C, int NT, int NCH>
hipError_t launch_chunk_nt(const StepArgs& a, hipStream_t s)
{
    const size_t shmem = sizeof(double) * (NT / 64) * C * NCH * 3 +
                         (sizeof(HT) + sizeof(int32_t)) * (size_t)a.chunk_ecap + sizeof(int32_t) * (a.chunk_tmax + 1);
    hipError_t e = hipFuncSetAttribute((const void*)kpm_step_chunk_kernel<HT, XT, C, NT, NCH>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((kpm_step_chunk_kernel<HT, XT, C, NT, NCH>), dim3(a.grid), dim3(NT), shmem, s, a);
    return hipGetLastError();
}

template <typename HT, typename XT, int C, int NCH>
hipError_t launch_chunk_r(const StepArgs& a, hipStream_t s)
{
    switch (a.chunk_threads) {
    case 256: return launch_chunk_nt<HT, XT, C, 256, NCH>(a, s);
    case 512: return launch_chunk_nt<HT, XT, C, 512, NCH>(a, s);
    case 1024: return launch_chunk_nt<HT, XT, C, 1024, NCH>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename HT, typename XT>
hipError_t launch_chunk(const StepArgs& a, hipStream_t s)
{
    constexpr int C = 128 / (int)sizeof(XT);   // one 128-byte line per row and chunk
    if (a.R == 64) return launch_chunk_r<HT, XT, C, 64 / C>(a, s);
    if (a.R == 128) return launch_chunk_r<HT, XT, C, 128 / C>(a, s);
    return hipErrorInvalidValue;
}

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

// ---- variant 17: several rows per wavefront with 16-byte gathers (R = 16, 32) ----
// The general form of variant 15 for smaller vector blocks: a row of T_j is
// R x element bytes, each lane carries 16 bytes of it (two complex FP32 or one
// complex FP64 vectors), so a wavefront covers RPW = 64 x 16 / (R x element)
// rows at once and one gather instruction fetches a neighbour row for each of
// them.  Each row's entries come through the scalar path in BATCH-entry chunks,
// the column chunks one chunk ahead; every lane group selects its own row's
// entry, and every row is summed entry by entry in CSR order.
template <typename T, int N> struct Chunks { T v[N]; };

template <typename HT, typename XT, int RPW, int BATCH>
__global__ __launch_bounds__(kThreads) void kpm_step_rows_kernel(const int64_t* __restrict__ rp,
                                                                 const int32_t* __restrict__ col,
                                                                 const HT* __restrict__ val,
                                                                 const void* __restrict__ cur_v,
                                                                 void* __restrict__ po_v,
                                                                 const void* __restrict__ vin_v, int nrows,
                                                                 double alpha, int beta, double* __restrict__ part)
{
    constexpr int VL = 16 / (int)sizeof(XT);   // vectors per lane
    constexpr int LPR = 64 / RPW;              // lanes per row
    constexpr int R = LPR * VL;
    using LT = typename std::conditional<VL == 2, float4, double2>::type;   // 16-byte lane element
    using AT = typename acc_of<XT>::type;
    const LT* __restrict__ cur = static_cast<const LT*>(cur_v);
    LT* __restrict__ po = static_cast<LT*>(po_v);
    const LT* __restrict__ vin = static_cast<const LT*>(vin_v);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int grp = lane / LPR;                // row slot of this lane
    const int gl = lane % LPR;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    const int stride = G * kWaves;
    const int nset = (nrows + RPW - 1) / RPW;  // row sets of RPW consecutive rows
    const int64_t nnz = rp[nrows];             // chunk loads start at or before nnz (BATCH padding follows)
    const bool direct = vin != nullptr;

    double d1[VL], d2[VL], d3[VL];
#pragma unroll
    for (int v = 0; v < VL; ++v) { d1[v] = 0.0; d2[v] = 0.0; d3[v] = 0.0; }

    // row r of set s: q[r] .. e[r] (empty past the last row)
    auto extents = [&](int set, int64_t* q, int64_t* e) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) { q[r] = 0; e[r] = 0; }
        if (set < nset) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int row = min(set * RPW + r, nrows);
                q[r] = rp[row];
                e[r] = set * RPW + r < nrows ? rp[row + 1] : q[r];
            }
        }
    };
    int set = __builtin_amdgcn_readfirstlane(lb * kWaves + wave);
    int64_t q[RPW], e[RPW];
    extents(set, q, e);
    Chunks<ColChunk<BATCH>, RPW> cc;
#pragma unroll
    for (int r = 0; r < RPW; ++r) cc.v[r] = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(q[r], nnz));
    for (; set < nset; set += stride) {
        int64_t nq[RPW], ne[RPW];
        extents(set + stride, nq, ne);
        int64_t span = 0, my_beg = q[0], my_end = e[0];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            span = max(span, e[r] - q[r]);
            if (grp == r) { my_beg = q[r]; my_end = e[r]; }
        }
        const int row = set * RPW + grp;
        const bool live = row < nrows;
        const size_t own = (size_t)row * LPR + gl;
        LT oc, op;
        zero16(oc);
        zero16(op);
        if (live) {
            oc = cur[own];
            if (beta) op = po[own];
        }
        AT acc[VL];
#pragma unroll
        for (int v = 0; v < VL; ++v) zero(acc[v]);
        for (int64_t k = 0; k < span; k += BATCH) {
            LT x[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                int c = cc.v[0].c[u];
#pragma unroll
                for (int r = 1; r < RPW; ++r) c = grp == r ? cc.v[r].c[u] : c;
                x[u] = cur[(size_t)c * LPR + gl];
            }
            Chunks<ValChunk<HT, BATCH>, RPW> hc;
#pragma unroll
            for (int r = 0; r < RPW; ++r)
                hc.v[r] = *reinterpret_cast<const ValChunk<HT, BATCH>*>(val + min(q[r] + k, nnz));
            const bool more = k + BATCH < span;
#pragma unroll
            for (int r = 0; r < RPW; ++r)
                cc.v[r] = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(more ? q[r] + k + BATCH : nq[r], nnz));
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                HT h = hc.v[0].h[u];
#pragma unroll
                for (int r = 1; r < RPW; ++r) h = grp == r ? hc.v[r].h[u] : h;
                if (my_beg + k + u < my_end) {
                    XT xv[VL];
                    split16(x[u], xv);
#pragma unroll
                    for (int v = 0; v < VL; ++v) mac(acc[v], h, xv[v]);
                }
            }
        }
        if (span == 0) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) cc.v[r] = *reinterpret_cast<const ColChunk<BATCH>*>(col + min(nq[r], nnz));
        }
        if (live) {
            XT cv[VL], pv[VL], ov[VL];
            split16(oc, cv);
            split16(op, pv);
#pragma unroll
            for (int v = 0; v < VL; ++v) {
                ov[v] = comb(alpha, acc[v], pv[v]);
                d1[v] += nrm2(cv[v]);
                d2[v] += redot(ov[v], cv[v]);
            }
            po[own] = join16(ov);
            if (direct) {
                XT vv[VL];
                split16(vin[own], vv);
#pragma unroll
                for (int v = 0; v < VL; ++v) d3[v] += redot(vv[v], cv[v]);
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) { q[r] = nq[r]; e[r] = ne[r]; }
    }

    // fold the row slots (fixed xor tree), then the wavefronts
#pragma unroll
    for (int m = LPR; m < 64; m <<= 1) {
#pragma unroll
        for (int v = 0; v < VL; ++v) {
            d1[v] += __shfl_xor(d1[v], m);
            d2[v] += __shfl_xor(d2[v], m);
            d3[v] += __shfl_xor(d3[v], m);
        }
    }
    __shared__ double red[kWaves][R][3];
    if (grp == 0) {
#pragma unroll
        for (int v = 0; v < VL; ++v) {
            red[wave][gl * VL + v][0] = d1[v];
            red[wave][gl * VL + v][1] = d2[v];
            red[wave][gl * VL + v][2] = d3[v];
        }
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

// ---- variant 12: row groups -------------------------------------------------------
// A wavefront owns kGroupRows consecutive rows (spatial order) at once.  Their
// column lists are merged ahead of time, so every distinct neighbour row is
// gathered once for the group and multiplied into all kGroupRows accumulators
// (values are zero where a row lacks the column).  On the 1.05 deg 8x8 cell
// this gathers 57 % of the rows the per-row kernels gather, for 2.3x the value
// bytes.  Columns and values come through the scalar path in BATCH-entry chunks.
template <typename HT, typename XT, int BATCH>
__global__ __launch_bounds__(kThreads) void kpm_step_group_kernel(const int64_t* __restrict__ grp,
                                                                  const int32_t* __restrict__ gcol,
                                                                  const HT* __restrict__ gval,
                                                                  const XT* __restrict__ cur, XT* __restrict__ po,
                                                                  const XT* __restrict__ vin, int n, int ng,
                                                                  double alpha, int beta, double* __restrict__ part)
{
    using AT = typename acc_of<XT>::type;
    constexpr int K = kGroupRows;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);
    const int stride = G * kWaves;
    const bool direct = vin != nullptr;

    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    for (int g = __builtin_amdgcn_readfirstlane(lb * kWaves + wave); g < ng; g += stride) {
        const int64_t q0 = grp[g], q1 = grp[g + 1];
        AT acc[K];
#pragma unroll
        for (int k = 0; k < K; ++k) zero(acc[k]);
        for (int64_t q = q0; q < q1; q += BATCH) {
            const ColChunk<BATCH> cc = *reinterpret_cast<const ColChunk<BATCH>*>(gcol + q);
            const ValChunk<HT, BATCH * K> vc = *reinterpret_cast<const ValChunk<HT, BATCH * K>*>(gval + q * K);
            XT x[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) x[u] = cur[(size_t)cc.c[u] * 64 + lane];
#pragma unroll
            for (int u = 0; u < BATCH; ++u)
                if (q + u < q1) {
#pragma unroll
                    for (int k = 0; k < K; ++k) mac(acc[k], vc.h[u * K + k], x[u]);
                }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int row = g * K + k;
            if (row >= n) break;
            const size_t idx = (size_t)row * 64 + lane;
            const XT cv = cur[idx];
            XT p;
            if (beta) p = po[idx];
            else zero(p);
            const XT o = comb(alpha, acc[k], p);
            po[idx] = o;
            d1 += nrm2(cv);
            d2 += redot(o, cv);
            if (direct) d3 += redot(vin[idx], cv);
        }
    }

    __shared__ double red[kWaves][64][3];
    red[wave][lane][0] = d1;
    red[wave][lane][1] = d2;
    red[wave][lane][2] = d3;
    __syncthreads();
    for (int t = threadIdx.x; t < 64 * 3; t += kThreads) {
        const int r = t / 3, k = t % 3;
        double sum = red[0][r][k];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) sum += red[w][r][k];
        part[((size_t)lb * 64 + r) * 3 + k] = sum;
    }
}

template <typename HT, typename XT, int NG>
hipError_t launch_tiled(const StepArgs& a, hipStream_t s)
{
    const size_t shmem = (size_t)kTileLdsBytes + sizeof(int32_t) * (kTileLdsBytes / (kTileGroup * sizeof(XT)));
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)kpm_step_tiled_kernel<HT, XT, NG>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((kpm_step_tiled_kernel<HT, XT, NG>), dim3(a.grid), dim3(kThreads), shmem, s, a);
    return hipGetLastError();
}

template <typename HT, typename XT, int LPR, int VPL>
hipError_t launch_typed(const StepArgs& a, hipStream_t s)
{
    if (a.variant == 5 && a.ntiles > 0) {
        if (a.R == 64) return launch_tiled_wave<HT, XT, 1>(a, s);
        if (a.R == 128) return launch_tiled_wave<HT, XT, 2>(a, s);
    }
    if (a.variant == 11 && a.ntiles > 0) {
        switch (a.R) {
        case 16: return launch_tiled_dpp<HT, XT, 1>(a, s);
        case 32: return launch_tiled_dpp<HT, XT, 2>(a, s);
        case 64: return launch_tiled_dpp<HT, XT, 4>(a, s);
        case 128: return launch_tiled_dpp<HT, XT, 8>(a, s);
        default: break;
        }
    }
    if (a.variant == 4 && a.ntiles > 0) {
        switch (a.R) {
        case 16: return launch_tiled_row<HT, XT, 1>(a, s);
        case 32: return launch_tiled_row<HT, XT, 2>(a, s);
        case 64: return launch_tiled_row<HT, XT, 4>(a, s);
        case 128: return launch_tiled_row<HT, XT, 8>(a, s);
        default: break;
        }
    }
    if (a.variant == 3 && a.ntiles > 0) {
        switch (a.R) {
        case 16: return launch_tiled<HT, XT, 1>(a, s);
        case 32: return launch_tiled<HT, XT, 2>(a, s);
        case 64: return launch_tiled<HT, XT, 4>(a, s);
        case 128: return launch_tiled<HT, XT, 8>(a, s);
        default: break;
        }
    }
    if constexpr (sizeof(XT) >= 8) {
        if (a.variant == 13 && a.chunk_wg != nullptr) return launch_chunk<HT, XT>(a, s);
    }
    if (a.variant == 14) return launch_xchunk<HT, XT>(a, s);
    if constexpr ((std::is_same<XT, float2>::value || std::is_same<XT, double2>::value) && sizeof(HT) <= 8) {
        if (a.variant == 17 && a.prev == a.out) {
            constexpr int VL = 16 / (int)sizeof(XT);
            const int rpw = 64 * VL / a.R;   // rows per wavefront
            constexpr int BT = sizeof(HT) >= 16 ? 4 : 8;
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(a.grid), dim3(kThreads), 0, s, a.row_ptr, a.col,
                                   static_cast<const HT*>(a.val), a.cur, a.out, a.v, a.n, a.alpha, a.beta, a.part);
                return hipGetLastError();
            };
            if (rpw == 2) return go(kpm_step_rows_kernel<HT, XT, 2, BT>);
            if (rpw == 4) return go(kpm_step_rows_kernel<HT, XT, 4, sizeof(HT) >= 8 ? 4 : 8>);
            return hipErrorInvalidValue;
        }
    }
    if constexpr (sizeof(XT) == 8 && sizeof(HT) <= 8) {
        if (a.variant == 15 || a.variant == 16) {
            // complex FP32 vectors only (float2), R = 64, T_{j-1} and T_{j+1} in one buffer
            if (a.R != 64 || a.prev != a.out || sizeof(XT) != 2 * sizeof(float) || !std::is_same<XT, float2>::value)
                return hipErrorInvalidValue;
            const float4* cur4 = static_cast<const float4*>(a.cur);
            float4* po4 = static_cast<float4*>(a.out);
            const float4* v4 = static_cast<const float4*>(a.v);
            if (a.variant == 15)
                hipLaunchKernelGGL((kpm_step_pair_kernel<HT, 8>), dim3(a.grid), dim3(kThreads), 0, s, a.row_ptr, a.col,
                                   static_cast<const HT*>(a.val), cur4, po4, v4, a.n, a.alpha, a.beta, a.part);
            else
                hipLaunchKernelGGL((kpm_step_pair_kernel<HT, 16>), dim3(a.grid), dim3(kThreads), 0, s, a.row_ptr,
                                   a.col, static_cast<const HT*>(a.val), cur4, po4, v4, a.n, a.alpha, a.beta, a.part);
            return hipGetLastError();
        }
    }
    if (LPR == 64 && VPL == 1 && a.variant == 12 && a.ngroups > 0 && a.prev == a.out) {
        constexpr int GB = sizeof(HT) >= 16 ? 4 : 8;   // entries per scalar chunk
        hipLaunchKernelGGL((kpm_step_group_kernel<HT, XT, GB>), dim3(a.grid), dim3(kThreads), 0, s, a.grp_rp,
                           a.grp_col, static_cast<const HT*>(a.grp_val), static_cast<const XT*>(a.cur),
                           static_cast<XT*>(a.out), static_cast<const XT*>(a.v), a.n, a.ngroups, a.alpha, a.beta,
                           a.part);
        return hipGetLastError();
    }
    if (LPR == 64 && (a.variant == 8 || a.variant == 9) && a.prev == a.out) {
        const XT* cur = static_cast<const XT*>(a.cur);
        XT* po = static_cast<XT*>(a.out);
        if (a.variant == 8)
            hipLaunchKernelGGL((kpm_step_smem_kernel<HT, XT, VPL, 8>), dim3(a.grid), dim3(kThreads), 0, s, a.row_ptr,
                               a.col, static_cast<const HT*>(a.val), cur, po, static_cast<const XT*>(a.v), a.n, a.alpha,
                               a.beta, a.part, a.xcd_rows);
        else
            hipLaunchKernelGGL((kpm_step_smem_kernel<HT, XT, VPL, 16>), dim3(a.grid), dim3(kThreads), 0, s, a.row_ptr,
                               a.col, static_cast<const HT*>(a.val), cur, po, static_cast<const XT*>(a.v), a.n, a.alpha,
                               a.beta, a.part, a.xcd_rows);
        return hipGetLastError();
    }
    if (LPR == 64 && a.variant == 6)
        hipLaunchKernelGGL((kpm_step_wave_kernel<HT, XT, VPL, 8, 1024>), dim3(a.grid), dim3(1024), 0, s, a);
    else if (LPR == 64 && a.variant == 7)
        hipLaunchKernelGGL((kpm_step_wave_kernel<HT, XT, VPL, 8, 512>), dim3(a.grid), dim3(512), 0, s, a);
    else if (LPR == 64 && a.variant == 10)
        hipLaunchKernelGGL((kpm_step_wave_kernel<HT, XT, VPL, 8, kThreads, true>), dim3(a.grid), dim3(kThreads), 0, s, a);
    else if (LPR == 64 && (a.variant == 1 || a.variant >= 8))
        hipLaunchKernelGGL((kpm_step_wave_kernel<HT, XT, VPL, 8>), dim3(a.grid), dim3(kThreads), 0, s, a);
    else if (LPR == 64 && a.variant == 2)
        hipLaunchKernelGGL((kpm_step_wave_kernel<HT, XT, VPL, 16>), dim3(a.grid), dim3(kThreads), 0, s, a);
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
