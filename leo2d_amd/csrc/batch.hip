// batch.hip -- one Chebyshev recursion over many independent Hamiltonians.
//
// Many small KPM jobs (commensurate cells of uneven size, each at several
// Bloch k-points, each with its own block of R random-phase vectors) are run
// as one problem: their CSR matrices are stacked block-diagonally and their
// vector blocks concatenated row-wise, so one launch advances every job by one
// Chebyshev step.  The reference runs each such job as a separate MPI job
// (Locality::rootChebSolve / workerChebSolve, src/locality.cpp:1271-3575),
// each calling computeDosKPM's recursion (src/locality.cpp:5945-6136) on its
// own matrix; here the per-job moments come out of one segmented reduction.
//
// Rows are cut into tiles that never cross a job boundary.  A workgroup
// sweeps tiles (XCD-aware grid stride) and writes one partial per tile and
// vector, so the per-job reduction is a fixed-order sum over that job's
// contiguous tile range: deterministic, no atomics.
//
// Two lane layouts for R < 64 vectors (both parity-tested):
//   mode 0 ("row groups"):  64/R rows per wavefront, R lanes per row, entries
//                           of a row walked sequentially by its lanes;
//   mode 1 ("entry slots"): one row per wavefront, 64/R entries of the row in
//                           flight at once (lanes e, e+R, ...), folded with
//                           cross-lane shuffles at the end of the row;
//   mode 2 ("row chunks", R = 16 or 32): 64/R rows per wavefront as in mode 0,
//                           but each 16-lane row of lanes loads 16 of its
//                           matrix row's entries in one access and passes them
//                           on with DPP row broadcasts, so 16 gathers are in
//                           flight per lane and the next 16 entries load while
//                           they land.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "kpm_common.h"

namespace leo2d {
namespace {

// a += the value held by lane (lane ^ off), per component
__device__ inline void xor_add(double2& a, int off)
{
    a.x += __shfl_xor(a.x, off);
    a.y += __shfl_xor(a.y, off);
}
__device__ inline void xor_add(float2& a, int off)
{
    a.x += __shfl_xor(a.x, off);
    a.y += __shfl_xor(a.y, off);
}
__device__ inline void xor_add(double& a, int off) { a += __shfl_xor(a, off); }
__device__ inline void xor_add(float& a, int off) { a += __shfl_xor(a, off); }

template <typename HT, typename XT, int R, int MODE>
__global__ __launch_bounds__(kThreads) void batch_step_kernel(BatchStepArgs a)
{
    using AT = typename acc_of<XT>::type;
    constexpr int NG = 64 / R;   // row groups (mode 0) or entry slots (mode 1) per wavefront
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int grp = lane / R;
    const int r = lane % R;
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const int32_t* __restrict__ col = a.col;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    XT* out = static_cast<XT*>(a.out);   // holds T_{j-1} on entry, T_{j+1} on exit
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;

    __shared__ double red[kWaves][R][2];

    for (int t = lb; t < a.ntiles; t += G) {
        const int row0 = a.tile_row0[t], row1 = a.tile_row0[t + 1];
        double d1 = 0.0, d2 = 0.0;
        if (MODE == 0) {
            for (int base = row0 + wave * NG; base < row1; base += kWaves * NG) {
                const int row = base + grp;
                if (row >= row1) continue;
                AT acc;
                zero(acc);
                int64_t q = rp[row];
                const int64_t e = rp[row + 1];
                for (; q + 4 <= e; q += 4) {
                    const int c0 = col[q], c1 = col[q + 1], c2 = col[q + 2], c3 = col[q + 3];
                    const HT h0 = val[q], h1 = val[q + 1], h2 = val[q + 2], h3 = val[q + 3];
                    const XT x0 = cur[(size_t)c0 * R + r];
                    const XT x1 = cur[(size_t)c1 * R + r];
                    const XT x2 = cur[(size_t)c2 * R + r];
                    const XT x3 = cur[(size_t)c3 * R + r];
                    mac(acc, h0, x0);
                    mac(acc, h1, x1);
                    mac(acc, h2, x2);
                    mac(acc, h3, x3);
                }
                for (; q < e; ++q) mac(acc, val[q], cur[(size_t)col[q] * R + r]);
                const size_t idx = (size_t)row * R + r;
                const XT cv = cur[idx];
                XT p;
                if (beta) p = out[idx];
                else zero(p);
                const XT o = comb(alpha, acc, p);
                out[idx] = o;
                d1 += nrm2(cv);
                d2 += redot(o, cv);
            }
        } else {
            for (int row = row0 + wave; row < row1; row += kWaves) {
                const int64_t q0 = rp[row], q1 = rp[row + 1];
                AT acc;
                zero(acc);
                int64_t q = q0 + grp;
                for (; q + NG < q1; q += 2 * NG) {
                    const int c0 = col[q], c1 = col[q + NG];
                    const HT h0 = val[q], h1 = val[q + NG];
                    const XT x0 = cur[(size_t)c0 * R + r];
                    const XT x1 = cur[(size_t)c1 * R + r];
                    mac(acc, h0, x0);
                    mac(acc, h1, x1);
                }
                if (q < q1) mac(acc, val[q], cur[(size_t)col[q] * R + r]);
                // sum the entry slots: every lane ends with the row total
#pragma unroll
                for (int off = R; off < 64; off <<= 1) xor_add(acc, off);
                if (grp == 0) {
                    const size_t idx = (size_t)row * R + r;
                    const XT cv = cur[idx];
                    XT p;
                    if (beta) p = out[idx];
                    else zero(p);
                    const XT o = comb(alpha, acc, p);
                    out[idx] = o;
                    d1 += nrm2(cv);
                    d2 += redot(o, cv);
                }
            }
        }
        // fold the row groups of the wavefront (mode 1: only group 0 is non-zero)
#pragma unroll
        for (int off = R; off < 64; off <<= 1) {
            d1 += __shfl_xor(d1, off);
            d2 += __shfl_xor(d2, off);
        }
        if (lane < R) {
            red[wave][r][0] = d1;
            red[wave][r][1] = d2;
        }
        __syncthreads();
        if (threadIdx.x < 2 * R) {
            const int rr = threadIdx.x % R, k = threadIdx.x / R;
            double s = red[0][rr][k];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) s += red[w][rr][k];
            a.part[((size_t)t * R + rr) * 2 + k] = s;
        }
        __syncthreads();
    }
}

template <typename HT, typename XT, int R>
__global__ __launch_bounds__(kThreads) void batch_step_chunk_kernel(BatchStepArgs a)
{
    static_assert(R == 16 || R == 32, "16-lane row slots");
    using AT = typename acc_of<XT>::type;
    constexpr int RPW = 64 / R;          // matrix rows per wavefront
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane / R;
    const int vi = lane % R;
    const int dl = lane & 15;            // lane within its 16-lane row
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (b & 7) * (G >> 3) + (b >> 3);

    const HT* __restrict__ val = static_cast<const HT*>(a.val);
    const int32_t* __restrict__ col = a.col;
    const int64_t* __restrict__ rp = a.row_ptr;
    const XT* __restrict__ cur = static_cast<const XT*>(a.cur);
    XT* out = static_cast<XT*>(a.out);   // holds T_{j-1} on entry, T_{j+1} on exit
    const double alpha = a.alpha;
    const bool beta = a.beta != 0;

    __shared__ double red[kWaves][R][2];

    for (int t = lb; t < a.ntiles; t += G) {
        const int row0 = a.tile_row0[t], row1 = a.tile_row0[t + 1];
        double d1 = 0.0, d2 = 0.0;
        for (int base = row0 + wave * RPW; base < row1; base += kWaves * RPW) {
            const int row = base + slot;
            const bool live = row < row1;
            int64_t kb = 0;
            int cnt = 0;
            XT own_c, own_p;
            zero(own_c);
            zero(own_p);
            if (live) {
                kb = rp[row];
                cnt = (int)(rp[row + 1] - kb);
                const size_t idx = (size_t)row * R + vi;
                own_c = cur[idx];
                if (beta) own_p = out[idx];
            }
            int mx = cnt;
#pragma unroll
            for (int off = R; off < 64; off <<= 1) mx = max(mx, __shfl_xor(mx, off));
            mx = __builtin_amdgcn_readfirstlane(mx);
            int mc = 0;
            HT mh;
            zero(mh);
            if (dl < cnt) {
                mc = col[kb + dl];
                mh = val[kb + dl];
            }
            AT acc;
            zero(acc);
            for (int k0 = 0; k0 < mx; k0 += 16) {
                int pc = 0;
                HT ph;
                zero(ph);
                const int q = k0 + 16 + dl;
                if (q < cnt) {
                    pc = col[kb + q];
                    ph = val[kb + q];
                }
                const int m = cnt - k0;
                XT xs[16];
                xchunk_gather<XT, R, 0>(xs, m, mc, cur, vi);
                xchunk_mac<HT, XT, AT, 0>(acc, m, mh, xs);
                mc = pc;
                mh = ph;
            }
            if (live) {
                const size_t idx = (size_t)row * R + vi;
                const XT o = comb(alpha, acc, own_p);
                out[idx] = o;
                d1 += nrm2(own_c);
                d2 += redot(o, own_c);
            }
        }
        // fold the row slots of the wavefront, then the wavefronts
#pragma unroll
        for (int off = R; off < 64; off <<= 1) {
            d1 += __shfl_xor(d1, off);
            d2 += __shfl_xor(d2, off);
        }
        if (lane < R) {
            red[wave][vi][0] = d1;
            red[wave][vi][1] = d2;
        }
        __syncthreads();
        if (threadIdx.x < 2 * R) {
            const int rr = threadIdx.x % R, k = threadIdx.x / R;
            double s = red[0][rr][k];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) s += red[w][rr][k];
            a.part[((size_t)t * R + rr) * 2 + k] = s;
        }
        __syncthreads();
    }
}

// mu[p][r][.] from the tiles of job p (doubling identity from the first step):
//   j = 0: mu_0 = <T0|T0>, mu_1 = <T1|T0>;  j >= 1: mu_2j = 2<Tj|Tj> - mu_0,
//   mu_2j+1 = 2<Tj+1|Tj> - mu_1
// One workgroup per job.  Thread t sums the partials (r, k) = t mod 2R over
// the tiles tt = first + u + s*NS (u = t / 2R, NS = kReduceThreads / 2R) with
// four running sums, so several loads are in flight; the strands are then
// added in order u = 0, 1, ...  The order is fixed, so the moments are
// deterministic.
constexpr int kReduceThreads = 256;
__global__ __launch_bounds__(kReduceThreads) void batch_reduce_kernel(const double* part, const int32_t* job_tile0, int R,
                                                                      int M, int j, double* mu)
{
    __shared__ double strand[kReduceThreads];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const int w = 2 * R;
    const int NS = kReduceThreads / w;
    const int rk = t % w, u = t / w;
    const int r = rk % R, k = rk / R;
    const int t0 = job_tile0[p], t1 = job_tile0[p + 1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (u < NS) {
        int tt = t0 + u;
        for (; tt + 3 * NS < t1; tt += 4 * NS) {
            s0 += part[((size_t)tt * R + r) * 2 + k];
            s1 += part[((size_t)(tt + NS) * R + r) * 2 + k];
            s2 += part[((size_t)(tt + 2 * NS) * R + r) * 2 + k];
            s3 += part[((size_t)(tt + 3 * NS) * R + r) * 2 + k];
        }
        for (; tt < t1; tt += NS) s0 += part[((size_t)tt * R + r) * 2 + k];
    }
    strand[t] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (t < w) {
        double s = 0.0;
        for (int q = 0; q < NS; ++q) s += strand[q * w + t];
        double* m = mu + ((size_t)p * R + r) * M;
        if (j == 0) m[k] = s;
        else m[2 * j + k] = 2.0 * s - m[k];
    }
}

// start vectors of every job: random phases (vector r of job p is global vector
// p * R + r, orbital i is the row's index inside its job) or, with targets,
// the unit vector on orbital targets[p * R + r] of job p
__global__ void batch_init_kernel(void* x, int vec_type, const int64_t* job_row0, int P, int64_t nrows, int R,
                                  uint64_t seed, const int32_t* targets)
{
    const size_t total = (size_t)nrows * R;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int64_t row = (int64_t)(t / R);
        const int r = (int)(t % R);
        int lo = 0, hi = P - 1;   // last job whose first row <= row
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (job_row0[mid] <= row) lo = mid;
            else hi = mid - 1;
        }
        const uint64_t i = (uint64_t)(row - job_row0[lo]);
        double sn, cs;
        if (targets) {
            cs = (int64_t)targets[(size_t)lo * R + r] == (int64_t)i ? 1.0 : 0.0;
            sn = 0.0;
        } else {
            const double phi = random_phase_angle(seed, (uint64_t)lo * R + r, i);
            sincos(phi, &sn, &cs);
        }
        if (vec_type == VEC_CPX_F64) static_cast<double2*>(x)[t] = make_double2(cs, sn);
        else if (vec_type == VEC_CPX_F32) static_cast<float2*>(x)[t] = make_float2((float)cs, (float)sn);
        else if (vec_type == VEC_REAL_F64) static_cast<double*>(x)[t] = cs;
        else static_cast<float*>(x)[t] = (float)cs;
    }
}

// append one job's CSR to the stacked matrix: columns shifted by its first row,
// row pointers by its first entry
__global__ void batch_offset_kernel(const int32_t* col_in, int64_t nnz, int32_t row_off, int32_t* col_out,
                                    const int64_t* rp_in, int n, int64_t nnz_off, int64_t* rp_out)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nnz; q += stride)
        col_out[q] = col_in[q] + row_off;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= (size_t)n; i += stride)
        rp_out[i] = rp_in[i] + nnz_off;
}

template <typename HT, typename XT, int R>
hipError_t launch_mode(const BatchStepArgs& a, hipStream_t s)
{
    if constexpr (R == 16 || R == 32) {
        if (a.mode == 2) {
            hipLaunchKernelGGL((batch_step_chunk_kernel<HT, XT, R>), dim3(a.grid), dim3(kThreads), 0, s, a);
            return hipGetLastError();
        }
    } else {
        if (a.mode == 2) return hipErrorInvalidValue;
    }
    if (a.mode == 1)
        hipLaunchKernelGGL((batch_step_kernel<HT, XT, R, 1>), dim3(a.grid), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL((batch_step_kernel<HT, XT, R, 0>), dim3(a.grid), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

template <typename HT, typename XT>
hipError_t launch_r(const BatchStepArgs& a, hipStream_t s)
{
    switch (a.R) {
    case 4: return launch_mode<HT, XT, 4>(a, s);
    case 8: return launch_mode<HT, XT, 8>(a, s);
    case 16: return launch_mode<HT, XT, 16>(a, s);
    case 32: return launch_mode<HT, XT, 32>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_batch_step(int mat_type, int vec_type, const BatchStepArgs& a, hipStream_t s)
{
    if (a.grid <= 0 || a.grid % 8 != 0 || a.ntiles <= 0) return hipErrorInvalidValue;
    if (mat_type == MAT_CPX_F64 && vec_type == VEC_CPX_F64) return launch_r<double2, double2>(a, s);
    if (mat_type == MAT_REAL_F64 && vec_type == VEC_CPX_F64) return launch_r<double, double2>(a, s);
    if (mat_type == MAT_CPX_F32 && vec_type == VEC_CPX_F32) return launch_r<float2, float2>(a, s);
    if (mat_type == MAT_REAL_F32 && vec_type == VEC_CPX_F32) return launch_r<float, float2>(a, s);
    // real H with real start vectors (target starts): real vector blocks, half the gathered bytes
    if (mat_type == MAT_REAL_F64 && vec_type == VEC_REAL_F64) return launch_r<double, double>(a, s);
    if (mat_type == MAT_REAL_F32 && vec_type == VEC_REAL_F32) return launch_r<float, float>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_batch_reduce(const double* part, const int32_t* job_tile0, int P, int R, int M, int j, double* mu,
                               hipStream_t s)
{
    if (P <= 0 || 2 * R > kReduceThreads) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(P), dim3(kReduceThreads), 0, s, part, job_tile0, R, M, j, mu);
    return hipGetLastError();
}

hipError_t launch_batch_init(void* x, int vec_type, const int64_t* job_row0, int P, int64_t nrows, int R, uint64_t seed,
                             const int32_t* targets, hipStream_t s)
{
    // real vectors hold unit-vector starts only (random phases are complex)
    const bool real = vec_type == VEC_REAL_F64 || vec_type == VEC_REAL_F32;
    if (!real && vec_type != VEC_CPX_F64 && vec_type != VEC_CPX_F32) return hipErrorInvalidValue;
    if (real && !targets) return hipErrorInvalidValue;
    const size_t total = (size_t)nrows * R;
    const int blocks = (int)std::min<size_t>(8192, (total + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(batch_init_kernel, dim3(std::max(blocks, 1)), dim3(kThreads), 0, s, x, vec_type, job_row0, P,
                       nrows, R, seed, targets);
    return hipGetLastError();
}

hipError_t launch_batch_offset(const int32_t* col_in, int64_t nnz, int32_t row_off, int32_t* col_out,
                               const int64_t* rp_in, int n, int64_t nnz_off, int64_t* rp_out, hipStream_t s)
{
    const size_t work = std::max<size_t>((size_t)nnz, (size_t)n + 1);
    const int blocks = (int)std::min<size_t>(4096, (work + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(batch_offset_kernel, dim3(std::max(blocks, 1)), dim3(kThreads), 0, s, col_in, nnz, row_off,
                       col_out, rp_in, n, nnz_off, rp_out);
    return hipGetLastError();
}

}  // namespace leo2d
