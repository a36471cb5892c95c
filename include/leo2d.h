/*
 * leo2d.h -- C-ABI of libleo2d, the MI355X-native drop-in for LEO2D's KPM
 * hot path (Hamiltonian assembly -> Chebyshev recursion -> Jackson DoS).
 *
 * The reference runs this path inside Locality::workerChebSolve
 * (src/locality.cpp:2843-3575): per MPI job it updates positions
 * (setConfigPositions, :3577), fills a CSR SpMatrix (generateRealH :3900 /
 * generateCpxH :4485), runs computeDosKPM / computeDosTraceKPM (:5945 /
 * :6138) on SpMatrix::vectorMultiply (src/matrix/spmatrix.cpp:270,365) and
 * transforms the moments with Param_tools::densityTransform{,Trace}
 * (src/params/param_tools.cpp:661,707).  The pair lists it assembles from are
 * built once by Locality::constructGeom (:785) through Hstruct::getInterPairs
 * (src/geom/hstruct.cpp:298) and Sheet::getIntraPairs (src/geom/sheet.cpp:624).
 * This header replaces that body; each entry point cites what it replaces.
 *
 * Conventions: plain pointers and sizes, no C++ types.  Every call returns
 * LEO2D_OK (0) or a negative LEO2D_ERR_* code; no exception crosses the ABI;
 * leo2d_last_error() gives the message.  A handle owns one HIP stream and its
 * device memory; host buffers belong to the caller.  A handle is not
 * thread-safe: use one handle per host thread.
 */
#ifndef LEO2D_H
#define LEO2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LEO2D_OK 0
#define LEO2D_ERR_ARG -1      /* invalid argument or size */
#define LEO2D_ERR_DECK -2     /* deck parse / semantic error */
#define LEO2D_ERR_HIP -3      /* HIP runtime error */
#define LEO2D_ERR_ALLOC -4    /* device or host allocation failed */
#define LEO2D_ERR_STATE -5    /* call out of order (e.g. KPM before a matrix) */
#define LEO2D_ERR_SPECTRUM -6 /* |mu_n| > mu_0 (1 + 1e-6): spectrum not inside [-1, 1] */
#define LEO2D_ERR_NAN -7      /* non-finite matrix value or moment */

#define LEO2D_MAX_SHEETS 8
#define LEO2D_MAX_TARGETS 64

typedef struct leo2d_handle leo2d_handle;
typedef struct leo2d_deck leo2d_deck;
typedef struct leo2d_geom leo2d_geom;

/* ---- deck: app/src/main.cpp:75-937 (same keys, same token rules) -------- */

/* Parse an hstruct deck given as text.  err (may be NULL) receives the message
 * of a LEO2D_ERR_DECK failure (e.g. POLY_ORDER % 4 != 0, main.cpp:914-917;
 * unknown MATERIAL, materials.cpp:28). */
int leo2d_deck_parse(const char* text, leo2d_deck** out, char* err, int errlen);
void leo2d_deck_free(leo2d_deck* d);
/* Typed access to the parsed Job_params bag (job_params.cpp:18-156 defaults). */
int leo2d_deck_get_int(const leo2d_deck* d, const char* key, int64_t* value);
int leo2d_deck_get_double(const leo2d_deck* d, const char* key, double* value);
int leo2d_deck_get_string(const leo2d_deck* d, const char* key, char* buf, int buflen);

/* ---- geometry: Locality::setupSupercell (locality.cpp:131-400), Sheet
 * enumeration (sheet.cpp:217-285), Hstruct global index and positions
 * (hstruct.cpp:69-218) -- the arrays constructGeom broadcasts
 * (locality.cpp:1059-1173). ---------------------------------------------- */
typedef struct {
    int32_t n;                                /* total orbitals */
    int32_t num_sheets;
    int32_t sheet_count[LEO2D_MAX_SHEETS];    /* orbitals per sheet */
    int32_t boundary_condition;
    double supercell[2][2];                   /* global supercell (opts "supercell") */
    double angle[LEO2D_MAX_SHEETS];           /* radians, after supercell setup */
} leo2d_geom_info;

int leo2d_geom_build(const leo2d_deck* d, leo2d_geom** out);
void leo2d_geom_free(leo2d_geom* g);
int leo2d_geom_get_info(const leo2d_geom* g, leo2d_geom_info* info);
/* index_to_grid int32[4n] (i, j, l, s), index_to_pos double[3n] */
int leo2d_geom_arrays(const leo2d_geom* g, int32_t* index_to_grid, double* index_to_pos);

/* ---- jobs: rootChebSolve job list (locality.cpp:1222-1388 SQ/LC shifts,
 * 1615-1652 K_TYPE 0 grid) and Hstruct::getTargetList (hstruct.cpp:1049). -- */
typedef struct {
    int32_t job_id;                           /* 1-based, the .cheb JOBID */
    int32_t max_jobs;
    double shift[LEO2D_MAX_SHEETS][3];        /* per-sheet shifts, lattice units */
    int32_t k_sampling;
    double k[2];                              /* Bloch vector (1/A) */
    double k_grid[2];
    double energy_rescale, energy_shift;      /* H~ = (H + shift I) / rescale */
    int32_t elec_on, mag_on;
    double E, B;
    int32_t complex_matrix;                   /* locality.cpp:2906-2908 */
    int32_t uniform_strain;
    double u[2][2];
    int32_t num_targets;
    int32_t targets[LEO2D_MAX_TARGETS];
    /* per-sheet strain (open boundaries only; no counterpart in the reference):
     * r -> (I + sheet_strain[s]) r on sheet s after the shifts, before the
     * uniform strain.  Deck key HETEROSTRAIN = eps nu sets uniaxial strain
     * [[eps, 0], [0, -nu eps]] on the last sheet. */
    int32_t sheet_strain_on;
    double sheet_strain[LEO2D_MAX_SHEETS][2][2];
    /* STRAIN_TYPE (main.cpp:595-614): 0 none; 1 SUPERCELL -- the Fourier
     * relaxation field given to leo2d_set_relaxation moves every orbital after
     * the shifts (locality.cpp:3657-3711); 6 DIRECT -- the same with the
     * reference's built-in coefficients (strain.cpp:756-876).  With either, the
     * interlayer terms take each orbital's local twist from its nearest
     * neighbour (locality.cpp:665-731, 4261-4269, 5045-5052). */
    int32_t strain_type;
} leo2d_job;

int leo2d_jobs_count(const leo2d_deck* d, const leo2d_geom* g, int32_t* count);
int leo2d_job_get(const leo2d_deck* d, const leo2d_geom* g, int32_t index, leo2d_job* job);

/* ---- device handle ------------------------------------------------------ */
int leo2d_create(int32_t device, leo2d_handle** out);
void leo2d_destroy(leo2d_handle* h);
const char* leo2d_last_error(const leo2d_handle* h);

/* Upload the geometry (grids, index and positions) to the device. */
int leo2d_set_geometry(leo2d_handle* h, const leo2d_geom* g);

/* GPU pair search replacing Hstruct::getInterPairs (hstruct.cpp:298-632) and
 * Sheet::getIntraPairs (sheet.cpp:624-765): same windows, same keep rules,
 * same per-row order (stable column sort for periodic decks), so the CSR
 * pattern is bit-identical to the reference's.  Counts returned (may be NULL). */
int leo2d_pairs_build(leo2d_handle* h, int64_t* n_intra, int64_t* n_inter, int64_t* n_slots);
/* How the last leo2d_pairs_build wrote the lists: 1 = one wavefront per row with
 * the row's lists staged in LDS, 0 = one thread per row through a scratch copy
 * (rows longer than the LDS lists hold, or LEO2D_PAIRS_STAGED=0).  No reference
 * counterpart: a diagnostic of this implementation. */
int leo2d_pairs_method(leo2d_handle* h, int32_t* staged);
/* Copy a pair list back: which 0 = intra, 1 = inter; ij int32[2P], sc int32[2P],
 * t double[P] (intra couplings, may be NULL). */
int leo2d_pairs_copy(leo2d_handle* h, int32_t which, int32_t* ij, int32_t* sc, double* t);

#define LEO2D_PATTERN_REFERENCE 0   /* every reference slot, explicit zeros kept */
#define LEO2D_PATTERN_PRUNED 1      /* slots outside the coupling support dropped */

/* Per-job Hamiltonian fill on the device, replacing generateRealH /
 * generateCpxH (locality.cpp:3900-4321, 4485-5158) including
 * setConfigPositions (locality.cpp:3577-3660, 3849-3857).  nnz (may be NULL)
 * receives the stored entries. */
int leo2d_assemble(leo2d_handle* h, const leo2d_job* job, int32_t pattern, int64_t* nnz);

/* Install a host CSR as the device Hamiltonian (H~, already rescaled):
 * row_ptr int64[n+1], col int32[nnz], val double[nnz] or double[2 nnz]. */
int leo2d_csr_upload(leo2d_handle* h, int32_t n, int64_t nnz, const int64_t* row_ptr, const int32_t* col,
                     const double* val, int32_t is_complex);
/* Shape of the device Hamiltonian. */
int leo2d_csr_info(leo2d_handle* h, int32_t* n, int64_t* nnz, int32_t* is_complex);
/* Copy the device Hamiltonian to host buffers sized from leo2d_csr_info. */
int leo2d_csr_copy(leo2d_handle* h, int64_t* row_ptr, int32_t* col, double* val);
/* Gershgorin interval [lo, hi] of the device Hamiltonian (FP64). */
int leo2d_spectral_bounds(leo2d_handle* h, double* lo, double* hi);

/* ---- KPM: computeDosKPM / computeDosTraceKPM (locality.cpp:5945-6346) ---- */
#define LEO2D_START_TARGETS 0       /* e_t per vector (local DoS) */
#define LEO2D_START_VECTORS 1       /* caller vectors, [R][n] real or complex */
#define LEO2D_START_RANDOM_PHASE 2  /* exp(i phi), phi = hash(seed, r, i) in [0, 2 pi) */

#define LEO2D_PREC_FP64 0           /* FP64 recursion */
#define LEO2D_PREC_MIXED 1          /* FP32 recursion and matrix, FP64 moment sums */

#define LEO2D_SCHEME_REFERENCE 0    /* direct mu_j for j < M/2, doubling for j >= M/4 */
#define LEO2D_SCHEME_DOUBLING 1     /* doubling identity from j = 1 (no start vector kept) */

typedef struct {
    int32_t M;              /* moments, multiple of 4 (POLY_ORDER) */
    int32_t R;              /* start vectors */
    int32_t start;          /* LEO2D_START_* */
    const int32_t* targets; /* [R] for LEO2D_START_TARGETS */
    const double* vectors;  /* [R][n] (x2 interleaved when vectors_complex) */
    int32_t vectors_complex;
    uint64_t seed;          /* LEO2D_START_RANDOM_PHASE */
    int32_t precision;      /* LEO2D_PREC_* */
    int32_t scheme;         /* LEO2D_SCHEME_* */
    int32_t block;          /* vectors per device block (power of two <= 128): 0 = auto (<= 64) */
    int32_t grid;           /* workgroups per step launch (multiple of 8): 0 = auto */
    int32_t time_steps;     /* 1: time every fused step launch with HIP events */
    int32_t natural_order;  /* 1: recurse in the reference row order; 0: spatial (Morton) order
                               when the matrix came from leo2d_assemble (moments are unchanged
                               up to summation order) */
} leo2d_kpm_params;

typedef struct {
    double recursion_ms;    /* HIP events around all steps and reductions */
    double step_ms_avg;     /* mean duration of one fused step launch */
    int64_t steps;          /* fused step launches */
    int64_t nnz;            /* stored entries of H */
    int32_t n, R, M;
    int32_t grid;
    int32_t kernel;   /* step-kernel variant launched for the last block (0 = lane per vector,
                         1 = wavefront per row, 14 = vector chunks split over the XCDs, 15 = two
                         FP32 rows per wavefront, 20 = SELL-4-sigma LDS tiles, 21 = the same with
                         a copy wavefront) */
    int32_t reserved;
} leo2d_timing;

/* mu: host double[R][M]. */
int leo2d_kpm_moments(leo2d_handle* h, const leo2d_kpm_params* p, double* mu);
/* Allocation-free variant of leo2d_kpm_moments.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Same parameters and output; every device buffer of the recursion (vector
   blocks, start-vector staging, partials, moments, targets) is carved out of a
   static device array of 64 MiB that each GPU holds from leo2d_create until the
   process ends, so no device memory is allocated or freed.  Runs the FP64
   recursion in the reference row order (natural_order is taken as 1,
   time_steps as 0; the timing record is left unchanged); the moments equal
   leo2d_kpm_moments' with natural_order = 1 bit for bit.  Limits:
   precision != LEO2D_PREC_FP64 -> LEO2D_ERR_ARG; a workspace above 64 MiB
   (about 2 n B 16 bytes for complex vectors, plus a third block for the
   reference scheme with caller vectors) -> LEO2D_ERR_ARG with the size in
   leo2d_last_error.  Calls on one GPU share its array: calls that overlap on
   one GPU (several threads or handles on one device) race. */
int leo2d_kpm_moments_noalloc(leo2d_handle* h, const leo2d_kpm_params* p, double* mu);
int leo2d_last_timing(const leo2d_handle* h, leo2d_timing* t);

/* ---- batched jobs: many small KPM jobs in one recursion -----------------
 * The reference farms every (cell, shift, k-point) job to its own MPI worker
 * (Locality::rootChebSolve / workerChebSolve, locality.cpp:1271-3575), each
 * running computeDosKPM's recursion (locality.cpp:5945-6136) on its own
 * matrix.  Here the jobs' Hamiltonians are stacked block-diagonally on the
 * device and advanced together: one launch per Chebyshev step for all jobs,
 * the whole recursion optionally captured once in a hipGraph and replayed.
 * Each job gets R start vectors: random phases (vector r of job p is global
 * vector p*R + r of the LEO2D_START_RANDOM_PHASE generator, orbital index
 * local to the job) or unit vectors on R of its orbitals (local DoS, as
 * computeDosKPM's targets), and its moments by the doubling identity from the
 * first step. */
int leo2d_batch_clear(leo2d_handle* h);
/* Append the current device Hamiltonian (from leo2d_assemble or
 * leo2d_csr_upload, in its own row order) as the next job; *job (may be NULL)
 * receives its index.  All jobs of a batch must be all real or all complex. */
int leo2d_batch_append(leo2d_handle* h, int32_t* job);
int leo2d_batch_info(const leo2d_handle* h, int32_t* jobs, int64_t* rows, int64_t* nnz);

#define LEO2D_BATCH_ROW_GROUPS 0    /* 64/R rows per wavefront */
#define LEO2D_BATCH_ENTRY_SLOTS 1   /* one row per wavefront, 64/R entries in flight */
#define LEO2D_BATCH_ROW_CHUNKS 2    /* R = 16, 32: 64/R rows per wavefront, 16-entry chunks broadcast in 16-lane rows */
#define LEO2D_BATCH_SELL 3          /* complex vectors, R = 16 or 32 (FP32: 32): SELL-4-sigma row tiles cut at job
                                       boundaries, their neighbour rows staged in LDS, one workgroup per CU
                                       (tile_rows and grid are chosen by the library) */

typedef struct {
    int32_t M;            /* moments per vector, multiple of 4 */
    int32_t R;            /* vectors per job: 4, 8, 16 or 32 */
    int32_t start;        /* LEO2D_START_RANDOM_PHASE, or LEO2D_START_TARGETS (local DoS of R orbitals per job) */
    const int32_t* targets;   /* [jobs][R] orbital indices inside each job, for LEO2D_START_TARGETS */
    uint64_t seed;
    int32_t precision;    /* LEO2D_PREC_FP64 or LEO2D_PREC_MIXED */
    int32_t tune;         /* 1: time the candidate launch shapes on this batch and keep the fastest */
    int32_t graph;        /* 1: capture the whole recursion in a hipGraph and replay it */
    int32_t mode;         /* LEO2D_BATCH_* when tune = 0 */
    int32_t tile_rows;    /* rows per tile when tune = 0 (0 = 128) */
    int32_t grid;         /* workgroups per step when tune = 0 (0 = auto, else a multiple of 8) */
} leo2d_batch_params;

typedef struct {
    double recursion_ms;  /* HIP events around the recursion (graph launch or eager steps) */
    double tune_ms;       /* autotuning time (0 when the cached choice was used) */
    int32_t mode, tile_rows, grid, graph;   /* launch shape used */
    int32_t candidates;   /* shapes timed by the autotuner */
    int32_t jobs, ntiles;
    int64_t rows, nnz;
} leo2d_batch_timing;

/* mu: host double[jobs][R][M]. */
int leo2d_batch_kpm(leo2d_handle* h, const leo2d_batch_params* p, double* mu);
int leo2d_batch_last_timing(const leo2d_handle* h, leo2d_batch_timing* t);

/* ---- DoS: Param_tools::densityTransform (param_tools.cpp:661-705) --------
 * rho[t][k] = scale * DCT-III(g . mu[t])_k / (pi sqrt(a^2 - (E_k - b)^2)),
 * E_k = b + a cos((k + 1/2) pi / M).  mu, E (may be NULL), rho: host. */
int leo2d_dos(leo2d_handle* h, const double* mu, int32_t T, int32_t M, double a, double b, double scale, double* E,
              double* rho);

/* ---- whole deck: rootChebSolve + workerChebSolve on one GPU --------------
 * Writes <out_prefix>.cheb (Param_tools::save, param_tools.cpp:19-285) and,
 * when MATRIX_SAVE > 0, <out_prefix>_matrix.dat (locality.cpp:4366-4383,
 * 5213-5228).  out_prefix NULL = the deck's JOB_NAME. */
int leo2d_run_deck(leo2d_handle* h, const char* deck_text, const char* out_prefix, int32_t pattern);

/* ---- learned surrogate: local stacking configuration -> local moments -----
 * No counterpart in the reference.  A multilayer perceptron (Fourier features
 * -> hidden GELU layers -> outputs) with bfloat16 weights and activations and
 * FP32 accumulation on the matrix cores, trained on the GPU (Adam, FP32 master
 * weights) on local moments that leo2d_batch_kpm computes.  A configuration
 * is (b1, b2, exx, eyy, exy, theta): the shift of sheet 2 in fractional
 * lattice units, the symmetric strain tensor and the twist in degrees. */
#define LEO2D_MLP_FEATURES 64
#define LEO2D_CONFIG_DIM 6

typedef struct leo2d_mlp leo2d_mlp;
typedef struct {
    int32_t in;        /* LEO2D_MLP_FEATURES */
    int32_t hidden;    /* units per hidden layer */
    int32_t layers;    /* hidden layers */
    int32_t out;       /* outputs (moments x sublattices) */
    uint64_t seed;     /* Glorot-uniform initialisation */
} leo2d_mlp_desc;

typedef struct {
    int32_t epochs, batch;
    float lr, beta1, beta2, eps;   /* Adam */
    uint64_t seed;                 /* shuffling */
} leo2d_mlp_train_params;

/* feat[n][64] from cfg[n][6] (host) */
int leo2d_mlp_features(const double* cfg, int64_t n, float* feat);
int leo2d_mlp_create(int32_t device, const leo2d_mlp_desc* d, leo2d_mlp** out);
void leo2d_mlp_destroy(leo2d_mlp* m);
const char* leo2d_mlp_last_error(const leo2d_mlp* m);
/* Flat parameter vector: per layer l, W_l[d_{l+1}][d_l] row-major then b_l[d_{l+1}]. */
int64_t leo2d_mlp_num_params(const leo2d_mlp* m);
int leo2d_mlp_get_params(leo2d_mlp* m, float* p);
int leo2d_mlp_set_params(leo2d_mlp* m, const float* p);   /* also resets the optimiser */
/* y[n][out] = network(x[n][in]) (host buffers) */
int leo2d_mlp_forward(leo2d_mlp* m, const float* x, int64_t n, float* y);
/* Mean squared error over the n x out outputs and its gradient (flat, as the parameters). */
int leo2d_mlp_gradients(leo2d_mlp* m, const float* x, const float* t, int32_t n, double* loss, float* grad);
/* Adam over shuffled mini-batches; epoch_loss[e] = mean squared error of epoch e (may be NULL). */
int leo2d_mlp_train(leo2d_mlp* m, const float* x, const float* t, int64_t n, const leo2d_mlp_train_params* p,
                    double* epoch_loss);
/* Trains count networks of one shape together on the same data: network k shuffles
 * with seeds[k] and takes exactly the updates leo2d_mlp_train with that seed would
 * (bit for bit), but each network's epoch is one captured graph replayed on its own
 * stream, so the networks' small launches overlap.  p->seed is not used.
 * losses (may be NULL): [count][epochs] mean squared error per epoch. */
int leo2d_mlp_train_ensemble(leo2d_mlp* const* nets, int32_t count, const float* x, const float* t, int64_t n,
                             const leo2d_mlp_train_params* p, const uint64_t* seeds, double* losses);
/* Ensemble of count networks of one shape on one device: mean[n][out] (may be NULL) and
 * spread[n][blocks]: for each of `blocks` equal slices of the outputs (one per
 * sublattice, say), the root mean square over the slice of the ensemble
 * standard deviation. */
int leo2d_mlp_ensemble(leo2d_mlp* const* nets, int32_t count, const float* x, int64_t n, int32_t blocks, float* mean,
                       float* spread);

/* ---- relaxation field (STRAIN_TYPE = SUPERCELL) ------------------------------
 * StrainCalc::loadFourierConfigFile (strain.cpp:258-284) without the file:
 * coeffs[3][max_k (max_k + 1) / 2 + 1] are the u_x, u_y, u_z Fourier
 * coefficients in the supercell basis, in the file's order.  Jobs with
 * strain_type = 1 assembled on this handle use them; max_k < 0 clears them. */
int leo2d_set_relaxation(leo2d_handle* h, int32_t max_k, const double* coeffs);
/* max_k is at most 64.  Read such a file (max_k, then the three rows, whitespace separated) into
 * coeffs (NULL to ask for the size): *max_k receives max_k and *count the
 * number of doubles, 3 (max_k (max_k + 1) / 2 + 1). */
int leo2d_read_relaxation(const char* path, int32_t* max_k, double* coeffs, int64_t* count);

/* ---- KPM conductivity (OBSERVABLE_TYPE = 1) -------------------------------- */
/* leo2d_assemble plus the velocity matrices dxH and dyH on H's pattern: slot
 * values (x_row + image shift - x_col) t and (y_row + image shift - y_col) t,
 * summed like H's, with t the real coupling or, for a complex H, the phased
 * one (generateRealH / generateCpxH, locality.cpp:4182-4183, 4299-4300,
 * 4939-4940, 5113-5114). */
int leo2d_assemble_velocity(leo2d_handle* h, const leo2d_job* job, int32_t pattern, int64_t* nnz);
/* Copy dxH's and dyH's values (same pattern and order as leo2d_csr_copy; nnz
 * doubles each for a real H, 2 nnz interleaved for a complex one).  Either
 * pointer may be NULL. */
int leo2d_velocity_copy(leo2d_handle* h, double* val_dx, double* val_dy);

typedef struct {
    int32_t M;                 /* moments per direction: C is M x M per target */
    int32_t R;                 /* targets */
    const int32_t* targets;    /* [R] orbitals (start vectors e_t) */
    int32_t block;             /* targets per device pass (0: min(R, 16)) */
} leo2d_cond_params;

typedef struct {
    double recursion_ms;       /* both Chebyshev vector sets */
    double gemm_ms;            /* the moment GEMMs */
    int32_t block;             /* targets per device pass used */
    int32_t reserved;
} leo2d_cond_timing;

/* Locality::computeCondKPM (locality.cpp:6348-6727) on the device matrices of
 * leo2d_assemble_velocity: C[t][n][m] = (dxH T_n(H) alpha_0) . (T_m(H) e_t),
 * alpha_0 = -dxH e_t.  Both vector sets are kept in HBM and C comes from one
 * FP64 GEMM per target.  (The reference leaves T_{M-1} of the beta set
 * unassigned; here it is the Chebyshev vector.) */
int leo2d_cond_moments(leo2d_handle* h, const leo2d_cond_params* p, double* C /*[R][M][M]*/);
/* Param_tools::conductivityTransform / matrixResponseTransform
 * (param_tools.cpp:770-905) for T blocks: out = 2D REDFT01 of 2 g_i g_j C_ij
 * with the Jackson g_i / sqrt(2). */
int leo2d_cond_transform(leo2d_handle* h, int32_t M, int32_t T, const double* C, double* out);
int leo2d_cond_last_timing(leo2d_handle* h, leo2d_cond_timing* t);

#ifdef __cplusplus
}
#endif
#endif /* LEO2D_H */
