// mlp.hip -- BF16 MFMA kernels of the learned local-moment surrogate.
//
// The surrogate is a small multilayer perceptron (64 Fourier features of a
// local stacking configuration -> hidden GELU layers -> local Chebyshev
// moments).  Weights and activations are bfloat16, every product accumulates
// in FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16), and the optimiser
// keeps FP32 master weights.  There is no counterpart in the reference; the
// moments it learns come from the KPM recursion of this library.
//
// One GEMM kernel serves every product of training and inference through
// strides: C[m][n] = sum_k A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak] and
// B(k, n) = B[k*sbk + n*sbn], so X W^T (forward), dZ^T X (weight gradient)
// and dZ W (input gradient) need no transposed copies.  The epilogue fuses
// bias, GELU, its derivative and the bf16 store of the next layer's operand.
//
// Tiling: a 256-thread workgroup owns a 64 x 64 tile of C; its 4 wavefronts
// each own 32 x 32 = 2 x 2 MFMA tiles of 16 x 16.  K advances in steps of 32
// through LDS (A tile 64 x 32, B tile 32 x 64, bf16), two MFMA k-steps each.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "mlp.h"

namespace leo2d {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kT = 256;
constexpr int kTM = 64, kTN = 64, kTK = 32;

__device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                       // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// GELU, tanh form: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2/pi)
__device__ inline float gelu(float z)
{
    const float c = 0.7978845608028654f;
    return 0.5f * z * (1.0f + tanhf(c * (z + 0.044715f * z * z * z)));
}
__device__ inline float gelu_grad(float z)
{
    const float c = 0.7978845608028654f;
    const float u = c * (z + 0.044715f * z * z * z);
    const float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

template <int EPI>
__global__ __launch_bounds__(kT) void gemm_bf16_kernel(GemmArgs g)
{
    __shared__ uint16_t As[kTM][kTK + 4];
    __shared__ uint16_t Bs[kTK][kTN + 4];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;

    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        // stage A (64 x 32) and B (32 x 64): 8 elements per thread each, zero outside the matrix
        for (int e = tid; e < kTM * kTK; e += kT) {
            const int mm = e / kTK, kk = e % kTK;
            const int m = m0 + mm, k = k0 + kk;
            As[mm][kk] = (m < g.M && k < g.K) ? g.A[(int64_t)m * g.sam + (int64_t)k * g.sak] : (uint16_t)0;
        }
        for (int e = tid; e < kTK * kTN; e += kT) {
            const int kk = e / kTN, nn = e % kTN;
            const int k = k0 + kk, n = n0 + nn;
            Bs[kk][nn] = (k < g.K && n < g.N) ? g.B[(int64_t)k * g.sbk + (int64_t)n * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kTK; ks += 16) {
            bf16x4 af[2], bfr[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int row = wm * 32 + i * 16 + (lane & 15);
#pragma unroll
                for (int t = 0; t < 4; ++t) af[i][t] = (short)As[row][ks + 4 * (lane >> 4) + t];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int cc = wn * 32 + j * 16 + (lane & 15);
#pragma unroll
                for (int t = 0; t < 4; ++t) bfr[j][t] = (short)Bs[ks + 4 * (lane >> 4) + t][cc];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    // epilogue: lane holds C[row0 + t][col] for t = 0..3
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn * 32 + j * 16 + (lane & 15);
            if (col >= g.N) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + t;
                if (row >= g.M) continue;
                const int64_t o = (int64_t)row * g.scm + col;
                const float v = acc[i][j][t];
                if (EPI == EPI_STORE_F32) {
                    g.C32[o] = v;
                } else if (EPI == EPI_BIAS_F32) {
                    g.C32[o] = v + g.bias[col];
                } else if (EPI == EPI_BIAS_GELU) {
                    const float z = v + g.bias[col];
                    g.Z[o] = z;
                    g.C16[o] = f2bf(gelu(z));
                } else {   // EPI_GELU_BWD: dZ = (dZ_next W) * gelu'(Z)
                    g.C16[o] = f2bf(v * gelu_grad(g.Z[o]));
                }
            }
        }
}

// ---- vectorised GEMM (gfx950 16x16x32 bf16 MFMA) -------------------------------
// Same product and epilogues.  Operands are staged as 16-byte runs of 8 bf16
// (along k when that is the contiguous direction, else along m / n), A as
// As[m][k] and B transposed as Bt[n][k] in LDS, so every MFMA operand is one
// 16-byte LDS read.  The next k-tile is loaded into registers while the
// current one is multiplied.  Needs K, the contiguous strides and the base
// addresses to be multiples of 8 elements (checked by the launcher).
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
constexpr int kPad = 8;

template <bool AK>
__device__ inline u16x8 load_a(const GemmArgs& g, int m0, int k0, int tid, int& mm, int& kk)
{
    // AK: 8 consecutive k of one row m (thread -> (m, k-run)); else 8 consecutive m of one k
    if (AK) {
        mm = tid >> 2;
        kk = (tid & 3) * 8;
    } else {
        kk = tid >> 3;
        mm = (tid & 7) * 8;
    }
    const int m = m0 + mm, k = k0 + kk;
    u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
    if (k < g.K && m < g.M) v = *reinterpret_cast<const u16x8*>(g.A + (int64_t)m * g.sam + (int64_t)k * g.sak);
    return v;
}
template <bool BK>
__device__ inline u16x8 load_b(const GemmArgs& g, int n0, int k0, int tid, int& nn, int& kk)
{
    if (BK) {
        nn = tid >> 2;
        kk = (tid & 3) * 8;
    } else {
        kk = tid >> 3;
        nn = (tid & 7) * 8;
    }
    const int n = n0 + nn, k = k0 + kk;
    u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
    if (k < g.K && n < g.N) v = *reinterpret_cast<const u16x8*>(g.B + (int64_t)k * g.sbk + (int64_t)n * g.sbn);
    return v;
}

template <int EPI, bool AK, bool BK>
__global__ __launch_bounds__(kT) void gemm_bf16_v2_kernel(GemmArgs g)
{
    __shared__ __align__(16) uint16_t As[kTM][kTK + kPad];
    __shared__ __align__(16) uint16_t Bt[kTN][kTK + kPad];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;

    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    int am, ak, bn, bk;
    u16x8 ra = load_a<AK>(g, m0, 0, tid, am, ak);
    u16x8 rb = load_b<BK>(g, n0, 0, tid, bn, bk);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        __syncthreads();   // previous tile fully read
        if (AK) {
            *reinterpret_cast<u16x8*>(&As[am][ak]) = ra;
        } else {
#pragma unroll
            for (int t = 0; t < 8; ++t) As[am + t][ak] = ra[t];
        }
        if (BK) {
            *reinterpret_cast<u16x8*>(&Bt[bn][bk]) = rb;
        } else {
#pragma unroll
            for (int t = 0; t < 8; ++t) Bt[bn + t][bk] = rb[t];
        }
        __syncthreads();
        if (k0 + kTK < g.K) {   // next tile in flight during the products
            ra = load_a<AK>(g, m0, k0 + kTK, tid, am, ak);
            rb = load_b<BK>(g, n0, k0 + kTK, tid, bn, bk);
        }
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            af[i] = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            bfr[j] = *reinterpret_cast<const bf16x8*>(&Bt[wn * 32 + j * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn * 32 + j * 16 + (lane & 15);
            if (col >= g.N) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + t;
                if (row >= g.M) continue;
                const int64_t o = (int64_t)row * g.scm + col;
                const float v = acc[i][j][t];
                if (EPI == EPI_STORE_F32) {
                    g.C32[o] = v;
                } else if (EPI == EPI_BIAS_F32) {
                    g.C32[o] = v + g.bias[col];
                } else if (EPI == EPI_BIAS_GELU) {
                    const float z = v + g.bias[col];
                    g.Z[o] = z;
                    g.C16[o] = f2bf(gelu(z));
                } else {
                    g.C16[o] = f2bf(v * gelu_grad(g.Z[o]));
                }
            }
        }
}

// dY = scale (Y - T) as bf16 and per-row squared-error sums (FP64, fixed order)
__global__ __launch_bounds__(kT) void mse_grad_kernel(const float* Y, const float* T, int rows, int cols, float scale,
                                                      uint16_t* dY, double* row_loss)
{
    const int r = blockIdx.x;
    if (r >= rows) return;
    __shared__ double s[kT];
    double acc = 0.0;
    for (int c = threadIdx.x; c < cols; c += kT) {
        const int64_t o = (int64_t)r * cols + c;
        const float d = Y[o] - T[o];
        dY[o] = f2bf(scale * d);
        acc += (double)d * (double)d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) row_loss[r] = s[0];
}

// db[n] = sum_m dZ[m][n] (bf16 operand values, FP32 sum in row order)
// db[c] = sum_r dZ[r][c].  A workgroup owns kColsumCols columns; its kColsumGroups
// row groups each sum a contiguous slice of rows (8 loads in flight), and the
// slices are added in group order, so the result does not depend on timing.
constexpr int kColsumCols = 16;
constexpr int kColsumGroups = kT / kColsumCols;

__global__ __launch_bounds__(kT) void colsum_kernel(const uint16_t* dZ, int rows, int cols, float* db)
{
    __shared__ float part[kColsumGroups][kColsumCols];
    const int cx = threadIdx.x % kColsumCols, g = threadIdx.x / kColsumCols;
    const int c = blockIdx.x * kColsumCols + cx;
    const int per = (rows + kColsumGroups - 1) / kColsumGroups;
    const int r0 = min(rows, g * per), r1 = min(rows, r0 + per);
    float s = 0.f;
    if (c < cols) {
        int r = r0;
        for (; r + 8 <= r1; r += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = bf2f(dZ[(int64_t)(r + u) * cols + c]);
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; r < r1; ++r) s += bf2f(dZ[(int64_t)r * cols + c]);
    }
    part[g][cx] = s;
    __syncthreads();
    if (g == 0 && c < cols) {
        float t = part[0][cx];
        for (int q = 1; q < kColsumGroups; ++q) t += part[q][cx];
        db[c] = t;
    }
}

// rows idx[b] of X (FP32) -> bf16 batch; rows idx[b] of T -> FP32 batch
__global__ __launch_bounds__(kT) void gather_kernel(const float* X, int in, const float* T, int out, const int64_t* idx,
                                                    int rows, uint16_t* Xb, float* Tb)
{
    const int b = blockIdx.x;
    if (b >= rows) return;
    const int64_t src = idx[b];
    for (int c = threadIdx.x; c < in; c += kT) Xb[(int64_t)b * in + c] = f2bf(X[src * in + c]);
    if (T)
        for (int c = threadIdx.x; c < out; c += kT) Tb[(int64_t)b * out + c] = T[src * out + c];
}

__global__ __launch_bounds__(kT) void cast_bf16_kernel(const float* x, int64_t n, uint16_t* y)
{
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) y[i] = f2bf(x[i]);
}

// Adam with bias correction; writes the bf16 copy used by the GEMMs
__global__ __launch_bounds__(kT) void adam_kernel(float* p, const float* grad, float* m, float* v, int64_t n, float lr,
                                                  float b1, float b2, float eps, float c1, float c2, uint16_t* p16)
{
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
        const float gr = grad[i];
        const float mi = b1 * m[i] + (1.f - b1) * gr;
        const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
        m[i] = mi;
        v[i] = vi;
        const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        const float pn = p[i] - upd;
        p[i] = pn;
        if (p16) p16[i] = f2bf(pn);
    }
}

// The same update with the bias corrections of step *ctr read from corr[2 *ctr],
// corr[2 *ctr + 1] (a captured training epoch replays unchanged while the step
// advances); count_kernel advances *ctr after it.
__global__ __launch_bounds__(kT) void adam_table_kernel(float* p, const float* grad, float* m, float* v, int64_t n,
                                                        float lr, float b1, float b2, float eps, const float* corr,
                                                        const int* ctr, uint16_t* p16)
{
    const int k = *ctr;
    const float c1 = corr[2 * (int64_t)k], c2 = corr[2 * (int64_t)k + 1];
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
        const float gr = grad[i];
        const float mi = b1 * m[i] + (1.f - b1) * gr;
        const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
        m[i] = mi;
        v[i] = vi;
        const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        const float pn = p[i] - upd;
        p[i] = pn;
        if (p16) p16[i] = f2bf(pn);
    }
}

__global__ void count_kernel(int* ctr)
{
    if (threadIdx.x == 0) *ctr += 1;
}

// per row: mean over the E ensemble members and, for each of nblk equal blocks
// of the outputs, the root mean square over the block of the ensemble standard
// deviation (block grid: blockIdx.y)
__global__ __launch_bounds__(kT) void ensemble_stats_kernel(const float* Ys, int E, int64_t estride, int rows, int cols,
                                                            int nblk, float* mean, float* spread)
{
    const int r = blockIdx.x, b = blockIdx.y;
    if (r >= rows || b >= nblk) return;
    const int bw = cols / nblk;
    __shared__ float s[kT];
    float acc = 0.f;
    for (int c = b * bw + threadIdx.x; c < (b + 1) * bw; c += kT) {
        const int64_t o = (int64_t)r * cols + c;
        float mu = 0.f;
        for (int e = 0; e < E; ++e) mu += Ys[e * estride + o];
        mu /= (float)E;
        float var = 0.f;
        for (int e = 0; e < E; ++e) {
            const float d = Ys[e * estride + o] - mu;
            var += d * d;
        }
        mean[o] = mu;
        acc += var / (float)E;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) spread[(int64_t)r * nblk + b] = sqrtf(s[0] / (float)bw);
}

int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + kT - 1) / kT)); }

}  // namespace

template <int EPI>
void launch_v2(const GemmArgs& g, bool ak, bool bk, dim3 grid, hipStream_t s)
{
    if (ak && bk) hipLaunchKernelGGL((gemm_bf16_v2_kernel<EPI, true, true>), grid, dim3(kT), 0, s, g);
    else if (ak) hipLaunchKernelGGL((gemm_bf16_v2_kernel<EPI, true, false>), grid, dim3(kT), 0, s, g);
    else if (bk) hipLaunchKernelGGL((gemm_bf16_v2_kernel<EPI, false, true>), grid, dim3(kT), 0, s, g);
    else hipLaunchKernelGGL((gemm_bf16_v2_kernel<EPI, false, false>), grid, dim3(kT), 0, s, g);
}

// the 16-byte staging needs: K a multiple of 8; a unit stride along k or along
// m (A) / n (B) with the other stride a multiple of 8; for m- / n-runs, M / N a
// multiple of 8; base addresses on 16 bytes
bool v2_ok(const GemmArgs& g, bool& ak, bool& bk)
{
    if (g.K % 8 != 0) return false;
    if (((uintptr_t)g.A | (uintptr_t)g.B) % 16 != 0) return false;
    if (g.sak == 1 && g.sam % 8 == 0) ak = true;
    else if (g.sam == 1 && g.sak % 8 == 0 && g.M % 8 == 0) ak = false;
    else return false;
    if (g.sbk == 1 && g.sbn % 8 == 0) bk = true;
    else if (g.sbn == 1 && g.sbk % 8 == 0 && g.N % 8 == 0) bk = false;
    else return false;
    return true;
}

hipError_t launch_gemm_bf16(const GemmArgs& g, int epi, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
    const dim3 grid((g.N + kTN - 1) / kTN, (g.M + kTM - 1) / kTM);
    bool ak = false, bk = false;
    if (v2_ok(g, ak, bk)) {
        switch (epi) {
        case EPI_STORE_F32: launch_v2<EPI_STORE_F32>(g, ak, bk, grid, s); break;
        case EPI_BIAS_F32: launch_v2<EPI_BIAS_F32>(g, ak, bk, grid, s); break;
        case EPI_BIAS_GELU: launch_v2<EPI_BIAS_GELU>(g, ak, bk, grid, s); break;
        case EPI_GELU_BWD: launch_v2<EPI_GELU_BWD>(g, ak, bk, grid, s); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (epi) {
    case EPI_STORE_F32: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_STORE_F32>, grid, dim3(kT), 0, s, g); break;
    case EPI_BIAS_F32: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_BIAS_F32>, grid, dim3(kT), 0, s, g); break;
    case EPI_BIAS_GELU: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_BIAS_GELU>, grid, dim3(kT), 0, s, g); break;
    case EPI_GELU_BWD: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_GELU_BWD>, grid, dim3(kT), 0, s, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_mse_grad(const float* Y, const float* T, int rows, int cols, float scale, uint16_t* dY,
                           double* row_loss, hipStream_t s)
{
    hipLaunchKernelGGL(mse_grad_kernel, dim3(rows), dim3(kT), 0, s, Y, T, rows, cols, scale, dY, row_loss);
    return hipGetLastError();
}

hipError_t launch_colsum(const uint16_t* dZ, int rows, int cols, float* db, hipStream_t s)
{
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + kColsumCols - 1) / kColsumCols), dim3(kT), 0, s, dZ, rows, cols, db);
    return hipGetLastError();
}

hipError_t launch_gather(const float* X, int in, const float* T, int out, const int64_t* idx, int rows, uint16_t* Xb,
                         float* Tb, hipStream_t s)
{
    hipLaunchKernelGGL(gather_kernel, dim3(rows), dim3(kT), 0, s, X, in, T, out, idx, rows, Xb, Tb);
    return hipGetLastError();
}

hipError_t launch_cast_bf16(const float* x, int64_t n, uint16_t* y, hipStream_t s)
{
    hipLaunchKernelGGL(cast_bf16_kernel, dim3(blocks_for(n)), dim3(kT), 0, s, x, n, y);
    return hipGetLastError();
}

hipError_t launch_ensemble_stats(const float* Ys, int E, int64_t estride, int rows, int cols, int nblk, float* mean,
                                 float* spread, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    if (nblk <= 0 || cols % nblk != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ensemble_stats_kernel, dim3(rows, nblk), dim3(kT), 0, s, Ys, E, estride, rows, cols, nblk, mean,
                       spread);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, int step, uint16_t* p16, hipStream_t s)
{
    float c1, c2;
    adam_corrections(b1, b2, step, &c1, &c2);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(kT), 0, s, p, grad, m, v, n, lr, b1, b2, eps, c1, c2,
                       p16);
    return hipGetLastError();
}

hipError_t adam_corrections(float b1, float b2, int step, float* c1, float* c2)
{
    *c1 = 1.f - powf(b1, (float)step);
    *c2 = 1.f - powf(b2, (float)step);
    return hipSuccess;
}

hipError_t launch_adam_table(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                             float eps, const float* corr, int* ctr, uint16_t* p16, hipStream_t s)
{
    hipLaunchKernelGGL(adam_table_kernel, dim3(blocks_for(n)), dim3(kT), 0, s, p, grad, m, v, n, lr, b1, b2, eps, corr,
                       ctr, p16);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(count_kernel, dim3(1), dim3(64), 0, s, ctr);
    return hipGetLastError();
}

}  // namespace leo2d
