/*
 * leo2d_oracle.h -- CPU FP64 restatement of LEO2D's KPM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (leo2d_amd) never links, imports or calls it.
 *
 * Every function restates one piece of the reference and cites it
 * (paths relative to the reference tree).  Floating-point expressions keep the
 * reference's operation order and are compiled without FMA contraction, so
 * positions, pair lists and real Hamiltonian values reproduce the reference's
 * arithmetic bit for bit on an x86-64 host.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   couplings      : against the reference's own graphene.cpp built into
 *                    oracle/_ref, plus closed-form known answers.
 *   sheet pairs    : against the reference's own Sheet::getIntraPairs built
 *                    into oracle/_ref (flake and periodic decks).
 *   inter pairs    : brute-force all-pairs scan restricted to the same window
 *                    rule (Hstruct is not buildable here: it needs fftw3.h).
 *   enumeration    : tests/tblg_test1.test target indices 60,61,175,176 and
 *                    the commensurate-cell count 4(N^2+NM+M^2).
 *   assembly       : Hermiticity, k=0 complex == real, dense consistency.
 *   KPM            : dense eigendecomposition of small H, mu0/mu1 closed
 *                    forms, doubling identity vs direct moments, and the
 *                    reference's SpMatrix::vectorMultiply from oracle/_ref.
 *   DoS            : O(M^2) direct DCT-III vs closed form, the Chebyshev-Gauss
 *                    normalisation invariant verified on the golden rows of
 *                    tests/tblg_test{1,2}.test.
 *   random phases  : no reference counterpart (new input form); its integer
 *                    hash is pinned against an independent Python restatement.
 *   Gaussian trace vectors: drawn by oracle.py's restatement of libstdc++'s
 *                    minstd_rand0 + normal_distribution, pinned bit for bit
 *                    against tests/golden/trace_gaussians.txt (written by a
 *                    libstdc++ program built here) and the standard's
 *                    10000th minstd_rand0 value 1043618065.
 */
#ifndef LEO2D_ORACLE_H
#define LEO2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SHEETS 8

/* materials */
#define ORC_MAT_GRAPHENE 0        /* src/materials/graphene.cpp */
#define ORC_MAT_GRAPHENE_SIMPLE 1 /* NN t=-2.7 eV, t_perp = 0.48 exp(-(r-3.35)/0.45), r<6 A */
#define ORC_MAT_GRAPHENE_SIMPLE_STRAIN 2 /* as 1, NN t = -2.7 exp(-3.37 (d/a_cc - 1)) at bond length d */
#define ORC_MAT_STRAINED_GRAPHENE 3 /* src/materials/strained_graphene.{h,cpp} */

typedef struct {
    int mat;
    int min_shape[2], max_shape[2];
    double max_R;
    int boundary_condition;
    double supercell[2][2];       /* sheet-local frame, rows = A1, A2 */
    int supercell_stride[2][2];
    double angle;                 /* radians */
    double height;
    int global_shifts_on;
    double global_shift[2];
} orc_sheet_desc;

typedef struct {
    int num_sheets;
    orc_sheet_desc sheet[ORC_MAX_SHEETS];
    int boundary_condition;
    int sc_search_size;
    double supercell[2][2];       /* global supercell (sheet 0's), opts "supercell" */
} orc_geom_desc;

typedef struct orc_geom orc_geom;

/* geometry ------------------------------------------------------------ */
orc_geom* orc_geom_build(const orc_geom_desc* d);
void orc_geom_free(orc_geom* g);
int orc_geom_n(const orc_geom* g);
int orc_geom_sheet_count(const orc_geom* g, int s);
void orc_geom_arrays(const orc_geom* g, int32_t* index_to_grid, double* index_to_pos);
int orc_grid_to_index(const orc_geom* g, int i, int j, int l, int s);
int orc_target_center(const orc_geom* g, int s);

/* couplings ------------------------------------------------------------ */
double orc_intralayer(int mat, int orb_row, int orb_col, int v0, int v1);
/* Coupling::Intralayer::strained_graphene with hop direction, strain and sheet
 * angle (strained_graphene.cpp:300-396) */
double orc_strained_intra(int orb_row, int orb_col, int v0, int v1, double hx, double hy, const double u[2][2],
                          double theta);
double orc_interlayer(int mat, int orb_row, int orb_col, double dx, double dy, double dz,
                      double theta_row, double theta_col);

/* pair lists: produced in reference order (per row: intra pairs of the
 * sheet, then inter pairs), ready for assembly. */
typedef struct orc_pairs orc_pairs;
orc_pairs* orc_pairs_build(const orc_geom* g);
void orc_pairs_free(orc_pairs* p);
int64_t orc_pairs_count(const orc_pairs* p, int which /*0 intra, 1 inter*/);
void orc_pairs_copy(const orc_pairs* p, int which, int32_t* ij /*[2P]*/, int32_t* sc /*[2P]*/,
                    double* t /*[P] (intra only, may be NULL)*/);

/* assembly -------------------------------------------------------------- */
typedef struct {
    double shift[ORC_MAX_SHEETS][2];   /* per sheet, fractional lattice units */
    double energy_rescale, energy_shift;
    int elec_on; double E;
    int mag_on;  double B;
    double k[2];
    int complex_matrix;
    int uniform_strain; double u[2][2];
    int pruned;                         /* 0 reference-exact slots, 1 drop zero slots */
    int sheet_strain_on;                /* per-sheet strain (open boundaries), before u */
    double sheet_strain[ORC_MAX_SHEETS][2][2];
    /* STRAIN_TYPE = SUPERCELL (1) or DIRECT (6): Fourier relaxation field and local
     * twist angles (6 has its coefficients built in) */
    int strain_type;
    int relax_max_k;
    const double* relax;                /* [3][relax_max_k (relax_max_k + 1) / 2 + 1] u_x, u_y, u_z */
    double relax_b[2][2];               /* reciprocal vectors of the supercell */
} orc_job;

/* StrainCalc::fourierStrainDisp_sc (strain.cpp:290-311, 390-460): relaxation
 * displacement of an orbital of sheet s at in-plane position r. */
void orc_relax_disp(const orc_job* job, double x, double y, int s, double* d /*[3]*/);
/* Locality::getLocalTheta (locality.cpp:665-731) on positions i2pos: the
 * sheet's local twist from an orbital's nearest neighbour, for every orbital. */
void orc_local_theta(const orc_geom* g, const double* i2pos, double* theta /*[n]*/);

typedef struct {
    int n;
    int64_t nnz;
    int64_t* row_ptr;   /* [n+1] */
    int32_t* col;       /* [nnz] */
    double* val;        /* [nnz] real or [2 nnz] complex interleaved */
    int is_complex;
    double* val_dx;     /* x-velocity matrix dxH on H's pattern, laid out like val */
    double* val_dy;     /* y-velocity matrix dyH, likewise */
} orc_csr;

void orc_job_positions(const orc_geom* g, const orc_job* job, double* i2pos /*[3n]*/);
int orc_assemble(const orc_geom* g, const orc_pairs* p, const orc_job* job, orc_csr* out);
void orc_csr_free(orc_csr* h);

/* KPM conductivity (Locality::computeCondKPM, locality.cpp:6348-6727; real H,
 * target starts): C[t][n*M + m] = (dxH T_n(H) alpha_0) . (T_m(H) e_t) with
 * alpha_0 = -dxH e_t.  The reference leaves T_{M-1} of the beta set
 * unassigned; here it is the Chebyshev vector itself. */
int orc_cond(int n, const int64_t* row_ptr, const int32_t* col, const double* val, const double* val_dx, int M,
             int T, const int32_t* targets, double* C /*[T][M*M]*/);
/* Param_tools::matrixResponseTransform (param_tools.cpp:838-905) for one
 * M x M block: out = 2D REDFT01 of 2 g_i g_j C_ij, g_i Jackson / sqrt(2). */
void orc_cond_transform(int M, const double* C, double* out);

/* KPM ------------------------------------------------------------------- */
#define ORC_START_TARGET 0      /* e_t, locality.cpp:5945 computeDosKPM */
#define ORC_START_VECTORS 1     /* given start vectors, locality.cpp:6138 computeDosTraceKPM */

int orc_kpm(int n, const int64_t* row_ptr, const int32_t* col, const double* val, int is_complex,
            int M, int R, int start, const int32_t* targets,
            const double* vecs, int vecs_complex, int doubling_all, int nthreads,
            double* mu /*[R][M]*/);

/* one faithful SpMV y = alpha*H*x + beta*y (spmatrix.cpp:346-360,395-410) */
void orc_spmv(int n, const int64_t* row_ptr, const int32_t* col, const double* val, int is_complex,
              const double* x, double* y, double alpha, double beta);

/* random-phase start vectors: v[r][i] = exp(i*phi(seed, r, i)) */
uint64_t orc_phase_hash(uint64_t seed, uint64_t r, uint64_t i);
void orc_random_phase(int n, int R, uint64_t seed, int r0, double* v /*[R][2n]*/);

/* DoS ------------------------------------------------------------------- */
void orc_jackson(int M, double* g);
void orc_energies(int M, double a, double b, double* E);
void orc_dos(int M, const double* mu, double a, double b, double* rho);
void orc_dos_trace(int M, int S, const double* mu, double a, double b, double scale, double* rho);

#ifdef __cplusplus
}
#endif
#endif
