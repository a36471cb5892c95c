// ref_harness.cpp -- C-ABI wrappers around the reference's own compiled sources
// (graphene.cpp, sheet.cpp, spmatrix.cpp), built by `make -C oracle ref` into
// oracle/_ref/libref.so.  TEST INFRASTRUCTURE ONLY: it lets tests pin the
// oracle's restatement against the reference's code on this host.  No
// reference source is copied here; this file only calls the reference's
// public functions.
#include <chrono>
#include <cstdio>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "geom/sdata.h"
#include "geom/sheet.h"
#include "materials/graphene.h"
#include "materials/materials.h"
#include "materials/strained_graphene.h"
#include "matrix/spmatrix.h"
#include "params/job_params.h"

// material of the sheets that ref_sheet_intra_pairs / ref_sheet_positions build
static Materials::Mat g_sheet_mat = Materials::Mat::Graphene;

extern "C" {

// 0: Graphene, 1: StrainedGraphene for the sheets built below
void ref_set_sheet_material(int strained)
{
    g_sheet_mat = strained ? Materials::Mat::StrainedGraphene : Materials::Mat::Graphene;
}

// Coupling::Intralayer::strained_graphene (strained_graphene.cpp:16), the unstrained table
double ref_strained_intra(int orb_row, int orb_col, int v0, int v1)
{
    std::array<int, 2> v = {{v0, v1}};
    return Coupling::Intralayer::strained_graphene(Graphene::orbital(orb_row), Graphene::orbital(orb_col), v);
}

// Coupling::Intralayer::strained_graphene (strained_graphene.cpp:300) with hop direction,
// strain u[4] = {u_xx, u_xy, u_yx, u_yy} and sheet angle
double ref_strained_intra_hop(int orb_row, int orb_col, int v0, int v1, double hx, double hy, const double* u,
                              double theta)
{
    std::array<int, 2> v = {{v0, v1}};
    std::vector<double> hop = {hx, hy};
    std::vector<std::vector<double> > strain = {{u[0], u[1]}, {u[2], u[3]}};
    return Coupling::Intralayer::strained_graphene(Graphene::orbital(orb_row), Graphene::orbital(orb_col), v, hop,
                                                   strain, theta);
}

// Coupling::Intralayer::graphene (graphene.cpp:13)
double ref_graphene_intra(int orb_row, int orb_col, int v0, int v1)
{
    std::array<int, 2> v = {{v0, v1}};
    return Coupling::Intralayer::graphene(Graphene::orbital(orb_row), Graphene::orbital(orb_col), v);
}

// Coupling::Interlayer::C_to_C (graphene.cpp:91)
double ref_graphene_inter(int orb_row, int orb_col, double x, double y, double z, double th_row, double th_col)
{
    std::array<double, 3> v = {{x, y, z}};
    return Coupling::Interlayer::C_to_C(Graphene::orbital(orb_row), Graphene::orbital(orb_col), v, th_row, th_col);
}

// Sheet::getIntraPairs (sheet.cpp:624) for one graphene sheet.  Returns the
// number of pairs; when the output pointers are non-null and cap suffices the
// pair list is copied out (ij[2P], sc[2P], t[P]).
long ref_sheet_intra_pairs(int min0, int min1, int max0, int max1, double max_R, int bc,
                           const double* supercell /*[4]*/, const int* stride /*[4]*/, int sc_search_size,
                           int global_shifts_on, const double* gshift /*[2]*/, int start_index,
                           int* ij, int* sc, double* t, long cap, int* n_orbitals_out)
{
    std::vector<int> mn = {min0, min1, min0};
    std::vector<int> mx = {max0, max1, max0};
    Job_params opts;
    opts.setParam("boundary_condition", bc);
    opts.setParam("sc_search_size", sc_search_size);
    opts.setParam("global_shifts_on", global_shifts_on);
    std::vector<std::vector<double> > gs(1, std::vector<double>(3, 0.0));
    gs[0][0] = gshift ? gshift[0] : 0.0;
    gs[0][1] = gshift ? gshift[1] : 0.0;
    opts.setParam("global_shifts", gs);
    Sdata sd(g_sheet_mat, mn, mx, max_R, bc, 0, 0, std::string(""), 0);
    sd.sheet_index = 0;
    sd.opts = opts;
    if (bc == 1) {
        sd.supercell = {{supercell[0], supercell[1]}, {supercell[2], supercell[3]}};
        sd.supercell_stride = {{stride[0], stride[1]}, {stride[2], stride[3]}};
    }
    Sheet sheet(sd);
    std::fflush(stdout);  // Sheet's constructor prints; keep it inside the caller's capture
    if (n_orbitals_out) *n_orbitals_out = sheet.getMaxIndex();
    std::vector<int> ai, aj;
    std::vector<double> at;
    std::vector<std::vector<int> > scv;
    sheet.getIntraPairs(ai, aj, at, scv, opts, start_index);
    long np = (long)ai.size();
    if (ij && np <= cap) {
        for (long q = 0; q < np; ++q) {
            ij[2 * q] = ai[q];
            ij[2 * q + 1] = aj[q];
            if (sc) { sc[2 * q] = scv[q][0]; sc[2 * q + 1] = scv[q][1]; }
            if (t) t[q] = at[q];
        }
    }
    return np;
}

// Sheet::posAtomIndex (sheet.cpp:458) for every orbital of the sheet above.
int ref_sheet_positions(int min0, int min1, int max0, int max1, double max_R, int bc,
                        const double* supercell, const int* stride, double* pos /*[3n]*/, int cap)
{
    std::vector<int> mn = {min0, min1, min0};
    std::vector<int> mx = {max0, max1, max0};
    Job_params opts;
    opts.setParam("boundary_condition", bc);
    Sdata sd(g_sheet_mat, mn, mx, max_R, bc, 0, 0, std::string(""), 0);
    sd.sheet_index = 0;
    sd.opts = opts;
    if (bc == 1) {
        sd.supercell = {{supercell[0], supercell[1]}, {supercell[2], supercell[3]}};
        sd.supercell_stride = {{stride[0], stride[1]}, {stride[2], stride[3]}};
    }
    Sheet sheet(sd);
    std::fflush(stdout);  // Sheet's constructor prints; keep it inside the caller's capture
    int n = sheet.getMaxIndex();
    if (pos && n <= cap)
        for (int k = 0; k < n; ++k)
            for (int d = 0; d < 3; ++d) pos[3 * k + d] = sheet.posAtomIndex(k, d);
    return n;
}

// SpMatrix::vectorMultiply (spmatrix.cpp:270, 365): y = alpha*H*x + beta*y.
void ref_spmv(int n, int nnz, const int* row_ptr, const int* col, const double* val, int is_complex,
              const double* x, double* y, double alpha, double beta)
{
    SpMatrix H;
    H.setup(nnz, n, n);
    int* ci = H.allocColIndx();
    int* rp = H.allocRowPtr();
    std::memcpy(rp, row_ptr, sizeof(int) * (n + 1));
    std::memcpy(ci, col, sizeof(int) * nnz);
    if (!is_complex) {
        double* v = H.allocRealVal();
        std::memcpy(v, val, sizeof(double) * nnz);
        H.vectorMultiply(const_cast<double*>(x), y, alpha, beta);
    } else {
        std::complex<double>* v = H.allocCpxVal();
        std::memcpy(v, val, sizeof(double) * 2 * nnz);
        H.vectorMultiply(reinterpret_cast<std::complex<double>*>(const_cast<double*>(x)),
                         reinterpret_cast<std::complex<double>*>(y), std::complex<double>(alpha),
                         std::complex<double>(beta));
    }
}

// Seconds spent in `reps` calls of SpMatrix::vectorMultiply (y <- H x, the
// reference's own CSR loop) on a matrix built once: the CPU timing of the
// reference's SpMV, with the matrix copy outside the timed calls.
double ref_spmv_time(int n, int nnz, const int* row_ptr, const int* col, const double* val, int is_complex,
                     const double* x, double* y, int reps)
{
    SpMatrix H;
    H.setup(nnz, n, n);
    int* ci = H.allocColIndx();
    int* rp = H.allocRowPtr();
    std::memcpy(rp, row_ptr, sizeof(int) * (n + 1));
    std::memcpy(ci, col, sizeof(int) * nnz);
    double* rv = nullptr;
    std::complex<double>* cv = nullptr;
    if (!is_complex) {
        rv = H.allocRealVal();
        std::memcpy(rv, val, sizeof(double) * nnz);
    } else {
        cv = H.allocCpxVal();
        std::memcpy(cv, val, sizeof(double) * 2 * nnz);
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) {
        if (!is_complex)
            H.vectorMultiply(const_cast<double*>(x), y, 1.0, 0.0);
        else
            H.vectorMultiply(reinterpret_cast<std::complex<double>*>(const_cast<double*>(x)),
                             reinterpret_cast<std::complex<double>*>(y), std::complex<double>(1.0),
                             std::complex<double>(0.0));
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The reference's KPM loop (Locality::computeDosKPM / computeDosTraceKPM,
// locality.cpp:5990-6041 and 6138-6346) restated around its own
// SpMatrix::vectorMultiply, for R start vectors on `nthreads` threads (vector r
// on thread r % nthreads; the matrix is built once and only read): per step one
// SpMV T_next = 2 H T_j, the subtract of T_prev and the two buffer copies, with
// mu_j = Re<v|T_j> for j < M/2 and the doubling identities mu_2j = 2<T_j|T_j> -
// mu_0, mu_2j+1 = 2 Re<T_j+1|T_j> - mu_1 for j >= M/4, as the reference fills
// cheb_coeffs.  vecs is [R][n] (real) or [R][2n] (complex, interleaved).
// Returns the seconds spent in the recursions (matrix copy outside).
double ref_kpm_time(int n, int nnz, const int* row_ptr, const int* col, const double* val, int is_complex,
                    const double* vecs, int R, int M, int nthreads, double* mu /*[R][M]*/)
{
    SpMatrix H;
    H.setup(nnz, n, n);
    int* ci = H.allocColIndx();
    int* rp = H.allocRowPtr();
    std::memcpy(rp, row_ptr, sizeof(int) * (n + 1));
    std::memcpy(ci, col, sizeof(int) * nnz);
    if (!is_complex) std::memcpy(H.allocRealVal(), val, sizeof(double) * nnz);
    else std::memcpy(H.allocCpxVal(), val, sizeof(double) * 2 * nnz);

    auto run_one = [&](int r) {
        double* m = mu + (size_t)r * M;
        if (is_complex) {
            typedef std::complex<double> cd;
            const cd* v = reinterpret_cast<const cd*>(vecs) + (size_t)r * n;
            std::vector<cd> T_prev(v, v + n), T_j(n), T_next(n);
            H.vectorMultiply(T_prev.data(), T_j.data(), cd(1.0), cd(0.0));
            double m0 = 0, m1 = 0;
            for (int c = 0; c < n; ++c) {
                m0 += std::real(std::conj(v[c]) * v[c]);
                m1 += std::real(std::conj(v[c]) * T_j[c]);
            }
            m[0] = m0;
            m[1] = m1;
            H.vectorMultiply(T_j.data(), T_next.data(), cd(2.0), cd(0.0));
            for (int c = 0; c < n; ++c) T_next[c] = T_next[c] - T_prev[c];
            for (int j = 2; j < M / 2; ++j) {
                for (int c = 0; c < n; ++c) {
                    T_prev[c] = T_j[c];
                    T_j[c] = T_next[c];
                }
                double mj = 0;
                for (int c = 0; c < n; ++c) mj += std::real(std::conj(v[c]) * T_j[c]);
                m[j] = mj;
                H.vectorMultiply(T_j.data(), T_next.data(), cd(2.0), cd(0.0));
                for (int c = 0; c < n; ++c) T_next[c] = T_next[c] - T_prev[c];
                if (j >= M / 4) {
                    double an_an = 0, anp_an = 0;
                    for (int c = 0; c < n; ++c) {
                        an_an += std::real(std::conj(T_j[c]) * T_j[c]);
                        anp_an += std::real(std::conj(T_next[c]) * T_j[c]);
                    }
                    m[2 * j] = 2 * an_an - m[0];
                    m[2 * j + 1] = 2 * anp_an - m[1];
                }
            }
        } else {
            const double* v = vecs + (size_t)r * n;
            std::vector<double> T_prev(v, v + n), T_j(n), T_next(n);
            H.vectorMultiply(T_prev.data(), T_j.data(), 1.0, 0.0);
            double m0 = 0, m1 = 0;
            for (int c = 0; c < n; ++c) {
                m0 += v[c] * v[c];
                m1 += v[c] * T_j[c];
            }
            m[0] = m0;
            m[1] = m1;
            H.vectorMultiply(T_j.data(), T_next.data(), 2.0, 0.0);
            for (int c = 0; c < n; ++c) T_next[c] = T_next[c] - T_prev[c];
            for (int j = 2; j < M / 2; ++j) {
                for (int c = 0; c < n; ++c) {
                    T_prev[c] = T_j[c];
                    T_j[c] = T_next[c];
                }
                double mj = 0;
                for (int c = 0; c < n; ++c) mj += v[c] * T_j[c];
                m[j] = mj;
                H.vectorMultiply(T_j.data(), T_next.data(), 2.0, 0.0);
                for (int c = 0; c < n; ++c) T_next[c] = T_next[c] - T_prev[c];
                if (j >= M / 4) {
                    double an_an = 0, anp_an = 0;
                    for (int c = 0; c < n; ++c) {
                        an_an += T_j[c] * T_j[c];
                        anp_an += T_next[c] * T_j[c];
                    }
                    m[2 * j] = 2 * an_an - m[0];
                    m[2 * j + 1] = 2 * anp_an - m[1];
                }
            }
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&, t]() {
            for (int r = t; r < R; r += nthreads) run_one(r);
        });
    for (auto& th : pool) th.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
