// cond.hip -- KPM conductivity moments (observable type 1) on the device.
//
// Replaces Locality::computeCondKPM (locality.cpp:6348-6727; real H, target
// starts) and Param_tools::matrixResponseTransform (param_tools.cpp:838-905).
// The reference builds beta_m = T_m(H) e_t (m < M), walks a_n = T_n(H) alpha_0
// with alpha_0 = -dxH e_t, and takes C[n][m] = (dxH a_n) . beta_m one dot
// product at a time: M^2 vector dots per target.  Here both vector sets are
// produced for a block of targets at once, one kernel per Chebyshev index
// writing straight into the [target][m][row] sets; C for every target is then
// an FP64 GEMM on the matrix cores (dgemm.hip) split into row chunks, and the 2D
// Jackson/DCT-III transform is two more GEMMs against the cosine matrix.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace leo2d {
namespace {

constexpr int kCondThreads = 256;

// set[(b * M + m) * n + i] = blk[i * B + b]: one Chebyshev vector block [n][B]
// stored as vector m of every target's set
__global__ __launch_bounds__(kCondThreads) void scatter_kernel(const double* __restrict__ blk, int n, int B, int m, int M,
                                                               double* __restrict__ set)
{
    const size_t total = (size_t)n * B;
    for (size_t t = blockIdx.x * (size_t)kCondThreads + threadIdx.x; t < total; t += (size_t)gridDim.x * kCondThreads) {
        const size_t i = t / B;
        const int b = (int)(t % B);
        set[((size_t)b * M + m) * n + i] = blk[t];
    }
}

// X[t][i][j] = 2 g_i g_j C[t][i][j] with the Jackson g_i / sqrt(2)
// (param_tools.cpp:857-890)
__global__ __launch_bounds__(kCondThreads) void cond_weight_kernel(const double* __restrict__ C, int M, int T,
                                                                   double* __restrict__ X)
{
    const size_t total = (size_t)T * M * M;
    const double pm = M_PI / M;
    for (size_t t = blockIdx.x * (size_t)kCondThreads + threadIdx.x; t < total; t += (size_t)gridDim.x * kCondThreads) {
        const int j = (int)(t % M);
        const int i = (int)((t / M) % M);
        const double gi = ((M - i) * cos(pm * i) + sin(pm * i) / tan(pm)) / (sqrt(2.0) * M);
        const double gj = ((M - j) * cos(pm * j) + sin(pm * j) / tan(pm)) / (sqrt(2.0) * M);
        X[t] = 2 * gi * gj * C[t];
    }
}

// column-major P(k, i) = w_i cos(pi i (k + 1/2) / M), w_0 = 1, w_i = 2 (FFTW
// REDFT01), the angle reduced exactly in integers: i (2k + 1) mod 4M
__global__ __launch_bounds__(kCondThreads) void dct3_matrix_kernel(int M, double* __restrict__ P)
{
    const size_t total = (size_t)M * M;
    const long long P4 = 4LL * M;
    for (size_t t = blockIdx.x * (size_t)kCondThreads + threadIdx.x; t < total; t += (size_t)gridDim.x * kCondThreads) {
        const int k = (int)(t % M);
        const int i = (int)(t / M);
        const long long m = ((long long)i * (2LL * k + 1)) % P4;
        const double c = cos(M_PI * (double)m / (2.0 * M));
        P[t] = i == 0 ? 1.0 : 2.0 * c;
    }
}

// C[b][q] = sum over the S row chunks c of part[b][c][q], in order c = 0, 1, ...
__global__ __launch_bounds__(kCondThreads) void chunk_sum_kernel(const double* __restrict__ part, int S, size_t mm,
                                                                 int B, double* __restrict__ C)
{
    const size_t total = (size_t)B * mm;
    for (size_t t = blockIdx.x * (size_t)kCondThreads + threadIdx.x; t < total; t += (size_t)gridDim.x * kCondThreads) {
        const size_t b = t / mm, q = t % mm;
        const double* p = part + b * S * mm + q;
        double s = 0.0;
        for (int c = 0; c < S; ++c) s += p[(size_t)c * mm];
        C[t] = s;
    }
}

// One Chebyshev index of the conductivity recursion for a block of B real
// vectors x = [n][B] (thread per (row, vector), rows in stored CSR order):
//   yh = alpha H x - (beta ? prev : 0)  -> out   [n][B] (the recursion; out may alias prev)
//   yd = dxH x                           -> dset[b][j][row] (written straight into the set)
// H and dxH share the pattern, so each entry's column and the gathered x are
// read once for both products.  With hx == nullptr only the dxH product is
// formed; with dx == nullptr only the H product, and hset (when given)
// receives yh as vector j of the set.
template <bool HAS_H, bool HAS_D>
__global__ __launch_bounds__(kCondThreads) void dual_step_kernel(const int64_t* __restrict__ rp,
                                                                 const int32_t* __restrict__ col,
                                                                 const double* __restrict__ hv,
                                                                 const double* __restrict__ dv, const double* x,
                                                                 const double* prev, double* out, double alpha, int beta,
                                                                 int n, int B, int j, int M, double* __restrict__ hset,
                                                                 double* __restrict__ dset)
{
    constexpr int U = 8;   // entries whose gathers are in flight together
    const size_t total = (size_t)n * B;
    for (size_t t = blockIdx.x * (size_t)kCondThreads + threadIdx.x; t < total; t += (size_t)gridDim.x * kCondThreads) {
        const int row = (int)(t / B);
        const int b = (int)(t % B);
        double ah = 0.0, ad = 0.0;
        int64_t q = rp[row];
        const int64_t e = rp[row + 1];
        for (; q + U <= e; q += U) {
            double xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[(size_t)col[q + u] * B + b];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (HAS_H) ah += hv[q + u] * xv[u];
                if (HAS_D) ad += dv[q + u] * xv[u];
            }
        }
        for (; q < e; ++q) {
            const double xv = x[(size_t)col[q] * B + b];
            if (HAS_H) ah += hv[q] * xv;
            if (HAS_D) ad += dv[q] * xv;
        }
        if (HAS_H) {
            const double y = alpha * ah - (beta ? prev[t] : 0.0);
            out[t] = y;
            if (hset) hset[((size_t)b * M + j) * n + row] = y;
        }
        if (HAS_D) dset[((size_t)b * M + j) * n + row] = ad;
    }
}

dim3 grid_for(size_t work) { return dim3((unsigned)std::min<size_t>(4096, (work + kCondThreads - 1) / kCondThreads)); }

}  // namespace

hipError_t launch_cond_scatter(const double* blk, int n, int B, int m, int M, double* set, hipStream_t s)
{
    hipLaunchKernelGGL(scatter_kernel, grid_for((size_t)n * B), dim3(kCondThreads), 0, s, blk, n, B, m, M, set);
    return hipGetLastError();
}

hipError_t launch_cond_chunk_sum(const double* part, int S, size_t mm, int B, double* C, hipStream_t s)
{
    hipLaunchKernelGGL(chunk_sum_kernel, grid_for((size_t)B * mm), dim3(kCondThreads), 0, s, part, S, mm, B, C);
    return hipGetLastError();
}

hipError_t launch_cond_dual_step(const int64_t* rp, const int32_t* col, const double* hv, const double* dv,
                                 const double* x, const double* prev, double* out, double alpha, int beta, int n, int B,
                                 int j, int M, double* hset, double* dset, hipStream_t s)
{
    const size_t work = (size_t)n * B;
    const dim3 g((unsigned)std::min<size_t>(8192, (work + kCondThreads - 1) / kCondThreads));
    if (hv && dv)
        hipLaunchKernelGGL((dual_step_kernel<true, true>), g, dim3(kCondThreads), 0, s, rp, col, hv, dv, x, prev, out,
                           alpha, beta, n, B, j, M, hset, dset);
    else if (hv)
        hipLaunchKernelGGL((dual_step_kernel<true, false>), g, dim3(kCondThreads), 0, s, rp, col, hv, dv, x, prev, out,
                           alpha, beta, n, B, j, M, hset, dset);
    else if (dv)
        hipLaunchKernelGGL((dual_step_kernel<false, true>), g, dim3(kCondThreads), 0, s, rp, col, hv, dv, x, prev, out,
                           alpha, beta, n, B, j, M, hset, dset);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_cond_weight(const double* C, int M, int T, double* X, hipStream_t s)
{
    hipLaunchKernelGGL(cond_weight_kernel, grid_for((size_t)T * M * M), dim3(kCondThreads), 0, s, C, M, T, X);
    return hipGetLastError();
}

hipError_t launch_dct3_matrix(int M, double* P, hipStream_t s)
{
    hipLaunchKernelGGL(dct3_matrix_kernel, grid_for((size_t)M * M), dim3(kCondThreads), 0, s, M, P);
    return hipGetLastError();
}

}  // namespace leo2d
