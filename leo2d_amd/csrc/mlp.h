// mlp.h -- launchers of the surrogate's BF16 MFMA kernels (mlp.hip).
// Internal to libleo2d; the public boundary is include/leo2d.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace leo2d {

// C[m][n] = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn]
// (bf16 operands stored as uint16_t, FP32 accumulation), then the epilogue:
enum GemmEpilogue {
    EPI_STORE_F32 = 0,   // C32[m*scm + n] = C
    EPI_BIAS_F32 = 1,    // C32 = C + bias[n]
    EPI_BIAS_GELU = 2,   // Z = C + bias[n] (FP32), C16 = bf16(gelu(Z))
    EPI_GELU_BWD = 3,    // C16 = bf16(C * gelu'(Z))
};
struct GemmArgs {
    const uint16_t* A;
    int64_t sam, sak;
    const uint16_t* B;
    int64_t sbk, sbn;
    int M, N, K;
    const float* bias;
    float* C32;
    uint16_t* C16;
    float* Z;
    int64_t scm;
};
hipError_t launch_gemm_bf16(const GemmArgs& g, int epi, hipStream_t s);

// dY = bf16(scale (Y - T)), row_loss[r] = sum_c (Y - T)^2
hipError_t launch_mse_grad(const float* Y, const float* T, int rows, int cols, float scale, uint16_t* dY,
                           double* row_loss, hipStream_t s);
// db[c] = sum_r dZ[r][c]
hipError_t launch_colsum(const uint16_t* dZ, int rows, int cols, float* db, hipStream_t s);
// Xb[b] = bf16(X[idx[b]]), Tb[b] = T[idx[b]] (T may be null)
hipError_t launch_gather(const float* X, int in, const float* T, int out, const int64_t* idx, int rows, uint16_t* Xb,
                         float* Tb, hipStream_t s);
hipError_t launch_cast_bf16(const float* x, int64_t n, uint16_t* y, hipStream_t s);
// mean[r][c] over E outputs Ys[e*estride + r*cols + c]; spread[r][b] = RMS over the
// outputs of block b (cols / nblk each) of the ensemble standard deviation
hipError_t launch_ensemble_stats(const float* Ys, int E, int64_t estride, int rows, int cols, int nblk, float* mean,
                                 float* spread, hipStream_t s);
// Adam step `step` (1-based) on FP32 master weights, refreshing their bf16 copy
hipError_t launch_adam(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, int step, uint16_t* p16, hipStream_t s);
// Adam's bias corrections 1 - b1^step, 1 - b2^step (host, the values launch_adam uses)
hipError_t adam_corrections(float b1, float b2, int step, float* c1, float* c2);
// Adam with the corrections of step *ctr taken from corr[2 *ctr .. 2 *ctr + 1], then *ctr += 1
hipError_t launch_adam_table(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                             float eps, const float* corr, int* ctr, uint16_t* p16, hipStream_t s);

}  // namespace leo2d
