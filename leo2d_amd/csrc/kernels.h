// kernels.h -- launchers for the HIP kernels (kpm.hip, dos.hip, assemble.hip).
// Internal to libleo2d; the public boundary is include/leo2d.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace leo2d {

// element types of a KPM launch
enum MatType { MAT_REAL_F64 = 0, MAT_CPX_F64 = 1, MAT_REAL_F32 = 2, MAT_CPX_F32 = 3 };
enum VecType { VEC_REAL_F64 = 0, VEC_CPX_F64 = 1, VEC_CPX_F32 = 2, VEC_REAL_F32 = 3 };

inline size_t vec_bytes(int vt) { return vt == VEC_CPX_F64 ? 16 : (vt == VEC_REAL_F32 ? 4 : 8); }

// One Chebyshev step over a block of R vectors stored row-major [n][R]:
//   out = alpha * H * cur - (beta ? prev : 0)        (out may alias prev)
// and per-vector partial sums for the moment reductions:
//   part[g][r][0] = sum_i |cur_ir|^2
//   part[g][r][1] = sum_i Re(conj(out_ir) cur_ir)
//   part[g][r][2] = sum_i Re(conj(v_ir) cur_ir)      (only when v != nullptr)
struct StepArgs {
    const int64_t* row_ptr;
    const int32_t* col;
    const void* val;
    int n;
    int R;
    const void* cur;
    const void* prev;
    void* out;
    const void* v;
    double alpha;
    int beta;
    double* part;   // [grid][R][3]
    int grid;       // number of workgroups (multiple of 8)
    int variant;    // 0 = lane per vector (any R), 1 = wavefront per row (R >= 64), 14 = vector chunks split
                    // over the XCDs (R = 16, 32, 64, 128), 15 = two rows per wavefront, 16-byte FP32 gathers
                    // (complex FP32, R = 64)
    int xcd_rows;   // variant 1: rows per XCD slice (0 = interleave rows over the grid)
};

// trailing (column 0, value 0) entries after nnz in the spatially ordered CSR
constexpr int kCsrPad = 16;
hipError_t launch_kpm_step(int mat_type, int vec_type, const StepArgs& a, hipStream_t s);

// ---- SELL-4-sigma step with LDS-staged row tiles (sell.hip) -----------------
// Tiles of at most kSellMaxRows consecutive rows whose distinct columns (the
// union, own rows first) fit `ucap` LDS rows of 256 B; each tile's rows sorted
// by length into groups of 4 (SELL-C-sigma with C = 4, sigma = the tile).
constexpr int kSellMaxRows = 64;    // rows per tile (8 wavefronts x 2 groups x 4 rows)
constexpr int kSellMaxLen = 64;     // longest row a group can hold
constexpr int kSellUcap = 288;      // staged rows per LDS buffer (two buffers of 288 x 256 B)
constexpr int kSellUcapTiled = 256; // the same with per-tile partials (their LDS takes the difference)
constexpr int kSellWaves = 8;       // wavefronts per workgroup (one workgroup per CU)
constexpr int kSellCwWaves = 7;     // compute wavefronts beside the copy wavefront (copy_wave)
constexpr int kSellCwRows = 56;     // rows per tile with a copy wavefront (7 wavefronts x 2 groups x 4 rows)
struct SellStepArgs {
    int n;
    int R;
    const void* cur;
    const void* prev;
    void* out;
    const void* v;
    double alpha;
    int beta;
    double* part;               // [grid][R][3], or [ntiles][R][2] with tile_part
    int grid;
    int tile_part;              // 1: partial sums per tile (stacked jobs), at most kSellUcapTiled union rows
    int ucap;                   // union rows the tiles were cut to
    int ntiles;
    const int32_t* tile_row0;   // [ntiles + 1]
    const int64_t* tile_u0;     // [ntiles + 1] offsets into ucols
    const int32_t* ucols;       // union rows of every tile (own rows first)
    const int32_t* tile_g0;     // [ntiles + 1] first group of each tile
    const int32_t* grow;        // [groups][4] row of each group slot (-1: empty)
    const int32_t* glen;        // [groups] steps of the group (multiple of 4, <= kSellMaxLen)
    const int64_t* goff;        // [groups] first entry of the group; entry e of slot s at goff + 4 e + s
    const uint16_t* eu;         // [entries] union index of the entry's column
    const void* ev;             // [entries] value (matrix element type)
    int tile_rows;              // most rows in a tile (<= kSellMaxRows)
    int max_len;                // most steps in a group (<= kSellMaxLen)
    int sliced;                 // 1: XCD x walks its own contiguous eighth of the tiles
    int reverse;                // 1: every workgroup walks its tiles last to first (alternate steps)
    int copy_wave;              // 1: staging copies issued by a wavefront of their own (tiles <= kSellCwRows)
};
// copy_wave: 16 complex FP32 vectors (128-byte row chunks) need the copy-wavefront kernel
bool sell_supported(int mat_type, int vec_type, int R, bool copy_wave);
hipError_t launch_kpm_step_sell(int mat_type, int vec_type, const SellStepArgs& a, hipStream_t s);
// dst[i] = src[map[i]] (zero where map[i] < 0), elements of elem_bytes (4, 8 or 16)
hipError_t launch_sell_fill(const int64_t* map, const void* src, void* dst, int elem_bytes, int64_t count,
                            hipStream_t s);

// Deterministic reduction of the partials of one step into moments.
//   scheme 0 (reference, target starts): mu_j = Re cur[target_r][r] for
//            2 <= j < M/2, dots used for j >= M/4 (locality.cpp:6002-6034)
//   scheme 1 (doubling from j = 1): mu_{2j}, mu_{2j+1} from the dots
//   scheme 2 (reference, vector starts): mu_j = part[..][2] for j < M/2
struct ReduceArgs {
    const double* part;
    int grid;
    int R;
    int M;
    int j;          // step index (0 = first product T1 = H T0)
    int scheme;
    const int32_t* targets;  // [R] (scheme 0)
    const void* cur;         // T_j   (scheme 0, j >= 2)
    const void* out;         // T_j+1 (scheme 0, j == 0)
    int vec_type;
    double* mu;              // [R][M] device
};
hipError_t launch_kpm_reduce(const ReduceArgs& a, hipStream_t s);

// ---- batched recursion over many stacked Hamiltonians (batch.hip) ----------
// One step for every job of a block-diagonal stack: out = alpha*H*cur - (beta ?
// out : 0) with per-tile partials part[t][r][0] = sum |cur|^2,
// part[t][r][1] = sum Re(conj(out) cur) over the rows of tile t.
struct BatchStepArgs {
    const int64_t* row_ptr;
    const int32_t* col;
    const void* val;
    const void* cur;
    void* out;                 // T_{j-1} in, T_{j+1} out
    double alpha;
    int beta;
    int R;                     // vectors per job: 4, 8, 16 or 32
    const int32_t* tile_row0;  // [ntiles + 1]; tiles never cross a job boundary
    int ntiles;
    double* part;              // [ntiles][R][2]
    int grid;                  // workgroups (multiple of 8)
    int mode;                  // 0 = row groups, 1 = entry slots, 2 = row chunks (R = 16, 32)
};
hipError_t launch_batch_step(int mat_type, int vec_type, const BatchStepArgs& a, hipStream_t s);
// mu[p][r][M] of job p from its tiles [job_tile0[p], job_tile0[p+1]) (doubling identity)
hipError_t launch_batch_reduce(const double* part, const int32_t* job_tile0, int P, int R, int M, int j, double* mu,
                               hipStream_t s);
// start vectors: random phases (vector r of job p = global vector p*R + r), or,
// when targets != nullptr, the unit vector on orbital targets[p*R + r] of job p
hipError_t launch_batch_init(void* x, int vec_type, const int64_t* job_row0, int P, int64_t nrows, int R, uint64_t seed,
                             const int32_t* targets, hipStream_t s);
hipError_t launch_batch_offset(const int32_t* col_in, int64_t nnz, int32_t row_off, int32_t* col_out,
                               const int64_t* rp_in, int n, int64_t nnz_off, int64_t* rp_out, hipStream_t s);

// start vectors
hipError_t launch_init_targets(void* x, int vec_type, int n, int R, const int32_t* targets, hipStream_t s);
hipError_t launch_init_random_phase(void* x, int vec_type, int n, int R, uint64_t seed, int r0, hipStream_t s);
hipError_t launch_init_from_host_layout(void* x, int vec_type, const double* src, int src_complex, int n, int R,
                                        hipStream_t s);
hipError_t launch_convert_f64_to_f32(const double* in, float* out, int64_t count, hipStream_t s);

// spatial permutation support (q = new row, i = iperm[q] = original row)
hipError_t launch_init_random_phase_perm(void* x, int vec_type, int n, int R, uint64_t seed, int r0,
                                        const int32_t* iperm, hipStream_t s);
hipError_t launch_init_from_host_layout_perm(void* x, int vec_type, const double* src, int src_complex, int n, int R,
                                             const int32_t* iperm, hipStream_t s);
hipError_t launch_perm_len(const int64_t* rp, const int32_t* iperm, int n, int64_t* len, hipStream_t s);
hipError_t launch_perm_fill(const int64_t* rp, const int32_t* col, const void* val, int val_bytes, const int32_t* iperm,
                            const int32_t* perm, int n, const int64_t* prp, int32_t* pcol, void* pval, hipStream_t s);
hipError_t scan_exclusive_i64(const int64_t* cnt, int64_t* out, int n, hipStream_t s);

// KPM conductivity (cond.hip): blk [n][B] -> set[b][m][n]; X = 2 g_i g_j C
// (Jackson g / sqrt 2); column-major DCT-III matrix P(k, i) = w_i cos(pi i (k+1/2)/M)
hipError_t launch_cond_scatter(const double* blk, int n, int B, int m, int M, double* set, hipStream_t s);
// one index of the conductivity recursion on a real [n][B] block: out = alpha H x - (beta ? prev : 0)
// (hset[b][j][.] = out when hset != nullptr) and dset[b][j][.] = dxH x; hv or dv may be nullptr
hipError_t launch_cond_dual_step(const int64_t* rp, const int32_t* col, const double* hv, const double* dv,
                                 const double* x, const double* prev, double* out, double alpha, int beta, int n, int B,
                                 int j, int M, double* hset, double* dset, hipStream_t s);
hipError_t launch_cond_weight(const double* C, int M, int T, double* X, hipStream_t s);
// C[b] = sum_c part[b][c] (c = 0 .. S-1, fixed order), blocks of mm values
hipError_t launch_cond_chunk_sum(const double* part, int S, size_t mm, int B, double* C, hipStream_t s);
hipError_t launch_dct3_matrix(int M, double* P, hipStream_t s);

// FP64 GEMM on the matrix cores (dgemm.hip), batched and split along k:
//   C[b][s](m, n) = sum_{k in split s} A_b(m, k) B_b(k, n)
//   A_b(m, k) = A[b sab + m sam + k sak], B_b(k, n) = B[b sbb + n sbn + k sbk],
//   C[b][s](m, n) at C[b scb + s scs + m scm + n scn]
// split s covers k in [s kc, (s+1) kc), kc = ceil(K / splits) rounded up to 16
struct DgemmArgs {
    int M, N, K;
    int batch;
    int splits;
    int kchunk;     // set by the launcher
    const double* A;
    int64_t sam, sak, sab;
    const double* B;
    int64_t sbn, sbk, sbb;
    double* C;
    int64_t scm, scn, scb, scs;
};
hipError_t launch_dgemm(const DgemmArgs& a, hipStream_t s);

// Jackson-kernel DoS: rho[t][k] = DCT-III(g * mu[t]) / (pi sqrt(a^2 - (E_k - b)^2)) * scale
hipError_t launch_dos(const double* mu, int T, int M, double a, double b, double scale, double* E, double* rho,
                      hipStream_t s);

// Static device workspace of the allocation-free KPM entry point (scratch.hip).
// scratch_init(d) looks up device d's address of the array (d must be current);
// scratch_ptr(d) returns it (nullptr before scratch_init succeeded).
constexpr size_t kScratchBytes = size_t(64) << 20;
hipError_t scratch_init(int device);
void* scratch_ptr(int device);
hipError_t launch_zero_f64(double* p, int64_t count, hipStream_t s);

// Gershgorin interval of a CSR matrix: lo = min_i (Re H_ii - R_i), hi = max_i (Re H_ii + R_i)
hipError_t launch_gershgorin(const int64_t* rp, const int32_t* col, const void* val, int is_complex, int n,
                             double* lohi /*[2*blocks]*/, int blocks, hipStream_t s);

}  // namespace leo2d
