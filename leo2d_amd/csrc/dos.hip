// dos.hip -- Jackson-kernel DoS reconstruction and the Gershgorin bound.
//
// DoS: Param_tools::densityTransform{,Trace} (param_tools.cpp:661-768) apply
// the Jackson damping g_n, an FFTW REDFT01 (DCT-III)
//     out_k = in_0 + 2 sum_{n>=1} in_n cos(pi n (k+1/2) / M)
// and divide by pi sqrt(a^2 - (E_k - b)^2).  Here one workgroup per moment row
// evaluates the DCT-III as a direct sum from an LDS cosine table indexed by the
// exact integer phase n(2k+1) mod 4M (table of 4M entries, 64 KiB at M = 2048),
// all in FP64.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace leo2d {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxLdsDoubles = 8192;   // 64 KiB of LDS per workgroup; larger tables live in global memory

__global__ __launch_bounds__(kThreads) void dos_kernel(const double* mu, int M, double a, double b, double scale,
                                                       double* E, double* rho, double* gtab)
{
    extern __shared__ double lds[];
    const int t = blockIdx.x;
    const int P = 4 * M;
    const bool in_lds = P + M <= kMaxLdsDoubles;
    double* cs = in_lds ? lds : gtab + (size_t)t * (P + M);
    double* in = in_lds ? lds + P : gtab + (size_t)t * (P + M) + P;
    const double pi = 3.14159265358979323846;
    for (int m = threadIdx.x; m < P; m += kThreads) cs[m] = cos(pi * (double)m / (2.0 * M));
    const double* mrow = mu + (size_t)t * M;
    for (int i = threadIdx.x; i < M; i += kThreads) {
        // Jackson coefficients (param_tools.cpp:670-671)
        const double g = ((M - i) * cos((pi * i) / M) + sin((pi * i) / M) / tan(pi / M)) / (M);
        in[i] = g * mrow[i];
        if (t == 0 && E) E[i] = b + a * cos((i * 1.0 + 0.5) * pi / M);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < M; k += kThreads) {
        const long long step = 2LL * k + 1;
        long long m = step % P;
        double s = 0.0;
        for (int q = 1; q < M; ++q) {
            s += in[q] * cs[m];
            m += step;
            if (m >= P) m %= P;
        }
        const double out = in[0] + 2.0 * s;
        const double Ek = b + a * cos((k * 1.0 + 0.5) * pi / M);
        rho[(size_t)t * M + k] = scale * out / (pi * sqrt(a * a - (Ek - b) * (Ek - b)));
    }
}

__global__ __launch_bounds__(kThreads) void gershgorin_kernel(const int64_t* rp, const int32_t* col, const void* val,
                                                              int is_complex, int n, double* lohi)
{
    __shared__ double slo[kThreads], shi[kThreads];
    double lo = 1e300, hi = -1e300;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        double d = 0.0, rsum = 0.0;
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
            double re_v, mag;
            if (is_complex) {
                const double2 h = static_cast<const double2*>(val)[q];
                re_v = h.x;
                mag = sqrt(h.x * h.x + h.y * h.y);
            } else {
                re_v = static_cast<const double*>(val)[q];
                mag = fabs(re_v);
            }
            if (col[q] == i) d += re_v;
            else rsum += mag;
        }
        lo = fmin(lo, d - rsum);
        hi = fmax(hi, d + rsum);
    }
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + w]);
            shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        lohi[2 * blockIdx.x] = slo[0];
        lohi[2 * blockIdx.x + 1] = shi[0];
    }
}

}  // namespace

hipError_t launch_dos(const double* mu, int T, int M, double a, double b, double scale, double* E, double* rho,
                      hipStream_t s)
{
    const int P = 4 * M;
    double* gtab = nullptr;
    size_t shmem = 0;
    if (P + M <= kMaxLdsDoubles) {
        shmem = sizeof(double) * (size_t)(P + M);
    } else {
        hipError_t e = hipMallocAsync((void**)&gtab, sizeof(double) * (size_t)T * (P + M), s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(dos_kernel, dim3(T), dim3(kThreads), shmem, s, mu, M, a, b, scale, E, rho, gtab);
    hipError_t e = hipGetLastError();
    if (gtab) hipFreeAsync(gtab, s);
    return e;
}

hipError_t launch_gershgorin(const int64_t* rp, const int32_t* col, const void* val, int is_complex, int n,
                             double* lohi, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(gershgorin_kernel, dim3(blocks), dim3(kThreads), 0, s, rp, col, val, is_complex, n, lohi);
    return hipGetLastError();
}

}  // namespace leo2d
