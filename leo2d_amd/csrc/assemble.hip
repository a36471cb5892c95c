// assemble.hip -- GPU pair search and per-job Hamiltonian fill.
//
// Pair search: every orbital row restates, candidate for candidate,
//   * Sheet::getIntraPairs (sheet.cpp:624-765): 8x8 unit-cell window around
//     the orbital's own cell in every periodic image; a candidate is kept when
//     its tabulated hopping is non-zero or it lies in the home image;
//   * Hstruct::getInterPairs (hstruct.cpp:298-632): 12x12 window (clamped to
//     the grid) on the sheets below and above; kept when the in-plane distance
//     is below 12.5 A;
// and, for periodic decks, the stable column sort of each block
// (Sheet::orderPairs sheet.cpp:810-849, Hstruct::orderPairs hstruct.cpp:634-666).
// Geometry arithmetic is compiled without FMA contraction and the rotation
// cosines come from the host's libm, so every keep/reject decision and window
// origin matches the reference bit for bit.  Two passes with a device scan in
// between: the counts, one thread per row; then the lists, one wavefront per row
// with the row's lists staged in LDS and the per-image sorted runs merged there
// (rows whose lists do not fit the LDS use the thread-per-row write pass, which
// merges through a scratch copy in global memory).
//
// Fill: one thread per row walks its pairs in the reference order and
// accumulates duplicate (row, col) pairs into one slot exactly as
// generateRealH / generateCpxH do (locality.cpp:3967-4316, 4664-5153):
// intralayer values from the hopping table, interlayer values from the
// coupling model at the job's positions, diagonal shift, rescale, Peierls and
// Bloch phases.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "assemble.h"
#include "kernels.h"

namespace leo2d {
namespace {

constexpr int kThreads = 256;
constexpr double kPI = 3.14159265358979323846;
constexpr double kPI_6 = 0.523598775598298873077;
constexpr double kA = 2.4768;
constexpr double kAcc = 1.429981146728865;   // a / sqrt(3)

inline dim3 rows_grid(int n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

// ---- couplings ----------------------------------------------------------
__device__ inline double intralayer(int mat, int orb_row, int orb_col, int v0, int v1)
{
    // Coupling::Intralayer::graphene (graphene.cpp:13-89): honeycomb shell index
    int h0 = 2 * v0 + v1;
    const int h1 = -1 * v0 + v1;
    if (orb_col == 0 && orb_row == 1) h0 -= 1;
    else if (orb_col == 1 && orb_row == 0) h0 += 1;
    const int r = h0 * (h0 + h1) + h1 * h1;
    if (mat == MAT_STRAINED_GRAPHENE) {
        // Coupling::Intralayer::strained_graphene, unstrained (strained_graphene.cpp:16-91)
        switch (r) {
        case 0: return 0.7833;
        case 1: return -2.8219;
        case 3: return 0.2543;
        case 4: return -0.1803;
        default: return 0.0;
        }
    }
    if (mat != MAT_GRAPHENE) return r == 1 ? -2.7 : 0.0;
    switch (r) {
    case 0: return 0.3504;
    case 1: return -2.8922;
    case 3: return 0.2425;
    case 4: return -0.2656;
    case 7: return 0.0235;
    case 9: return 0.0524;
    case 12: return -0.0209;
    case 13: return -0.0148;
    case 16: return -0.0211;
    default: return 0.0;
    }
}

// Coupling::Intralayer::strained_graphene with hop direction (hx, hy), strain u and
// sheet angle cos/sin (strained_graphene.cpp:300-396).  The reference writes its
// vector-like strain pair's first slot twice, so the pair is (u_xy + u_yx, 0).
__device__ inline double strained_intra(int orb_row, int orb_col, int v0, int v1, double hx, double hy,
                                        const double (&u)[2][2], double c, double s)
{
    int h0 = 2 * v0 + v1;
    const int h1 = -1 * v0 + v1;
    if (orb_col == 0 && orb_row == 1) h0 -= 1;
    else if (orb_col == 1 && orb_row == 0) h0 += 1;
    const int r = h0 * (h0 + h1) + h1 * h1;
    const double onedim = u[0][0] + u[1][1];
    const double two0 = u[0][1] + u[1][0], two1 = 0.0;
    const double w0 = c * hx + s * hy;
    const double w1 = -s * hx + c * hy;
    // unperturbed hopping t, scalar-strain a and vector-strain b of shells 0, 1, 3, 4
    // (strained_graphene.cpp:81-88)
    switch (r) {
    case 0: return 0.7833 + -4.8782 * onedim;
    case 1: return -2.8219 + 4.0066 * onedim + -3.0868 * (w0 * two0 + w1 * two1);
    case 3: return 0.2543 + -0.4633 * onedim + 0.8017 * (-w1 * two0 + w0 * two1);
    case 4: return -0.1803 + 0.6236 * onedim + 0.4793 * (w0 * two0 + w1 * two1);
    default: return 0.0;
    }
}

__device__ inline int iabs_trunc(double z)
{
    // graphene.cpp:117 applies the C library's int abs() to a double
    const int t = (int)z;
    return t < 0 ? -t : t;
}

__device__ double c_to_c(int orb_row, int orb_col, double v0, double v1, double v2, double theta_row,
                         double theta_col)
{
    // Coupling::Interlayer::C_to_C (graphene.cpp:91-222)
    if (iabs_trunc(v2) > 5) return 0.0;
    const double r2 = v0 * v0 + v1 * v1;
    if (!(r2 < 64.0)) return 0.0;   // sqrt(r2) < 8 exactly when r2 < 64
    const double r = sqrt(r2);
    const double ac = atan2(v1, v0);
    double theta21 = ac - theta_row;
    if (orb_row == 1) theta21 += kPI_6;
    else theta21 -= kPI_6;
    double theta12 = ac - theta_col + kPI;
    if (orb_col == 1) theta12 += kPI_6;
    else theta12 -= kPI_6;
    const double rs = r / kA;
    const double z_eq = 3.35;
    const double z_eps = (fabs(v2) - z_eq) / z_eq;
    const double lambda_0 = 0.310 - 1.882 * z_eps + 7.741 * z_eps * z_eps;
    const double xi_0 = 1.750 - 1.618 * z_eps + 1.848 * z_eps * z_eps;
    const double kappa_0 = 1.990 + 1.007 * z_eps + 2.427 * z_eps * z_eps;
    const double lambda_3 = -0.068 + 0.399 * z_eps - 1.739 * z_eps * z_eps;
    const double xi_3 = 3.286 - 0.914 * z_eps + 12.011 * z_eps * z_eps;
    const double x_3 = 0.500 + 0.322 * z_eps + 0.908 * z_eps * z_eps;
    const double lambda_6 = -0.008 + 0.046 * z_eps - 0.183 * z_eps * z_eps;
    const double xi_6 = 2.272 - 0.721 * z_eps - 4.414 * z_eps * z_eps;
    const double x_6 = 1.217 + 0.027 * z_eps - 0.658 * z_eps * z_eps;
    const double kappa_6 = 1.562 - 0.371 * z_eps - 0.134 * z_eps * z_eps;
    const double V0 = lambda_0 * exp(-xi_0 * (rs) * (rs)) * cos(kappa_0 * rs);
    const double V3 = lambda_3 * rs * rs * exp(-xi_3 * (rs - x_3) * (rs - x_3));
    const double V6 = lambda_6 * exp(-xi_6 * (rs - x_6) * (rs - x_6)) * sin(kappa_6 * rs);
    double t = V0 + V3 * (cos(3 * theta12) + cos(3 * theta21)) + V6 * (cos(6 * theta12) + cos(6 * theta21));
    const double d = 8.0 - r;
    if (d < 1. && d > 0.) t *= exp(1. - 1. / (d * d));
    return t;
}

__device__ inline double interlayer(int mat, int o1, int o2, double dx, double dy, double dz, double th1, double th2)
{
    if (mat != MAT_GRAPHENE && mat != MAT_STRAINED_GRAPHENE) {
        // t_perp(r) = 0.48 exp(-(r - 3.35)/0.45) eV for 3D distance r < 6 A
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < 36.0)) return 0.0;
        const double r = sqrt(r2);
        return 0.48 * exp(-(r - 3.35) / 0.45);
    }
    return c_to_c(o1, o2, dx, dy, dz, th1, th2);
}

// support of a coupling (pruned pattern): decided by exact comparisons only
__device__ inline bool inter_support(int mat, double dx, double dy, double dz)
{
    if (mat != MAT_GRAPHENE && mat != MAT_STRAINED_GRAPHENE) return dx * dx + dy * dy + dz * dz < 36.0;
    if (iabs_trunc(dz) > 5) return false;
    return dx * dx + dy * dy < 64.0;
}

// ---- pair search ----------------------------------------------------------
__device__ inline int grid_at(const SheetDev& s, int norb, int i, int j, int l)
{
    if (i >= 0 && i < s.H && j >= 0 && j < s.W && l >= 0 && l < norb) return s.grid[((size_t)i * s.W + j) * norb + l];
    return -1;
}

__device__ inline int clamp_cell(double v, int lo, int hi)
{
    int f = (int)floor(v);
    if (f < lo) f = lo;
    if (f > hi) f = hi;
    return f - lo;
}

// Each supercell image of the search yields its candidates in increasing column
// order (the windows walk i, j, l ascending, and the sheet below precedes the
// sheet above), so a list is a run of sorted runs, one per image; run_end
// records where each ends so the write pass can merge instead of sorting.
constexpr int kMaxRuns = 49;   // (2 S + 1)^2 images for SUPERCELL_SEARCH_SIZE S <= 3

struct PairSink {
    int32_t* col;
    int32_t* sc;   // (dx & 0xffff) | (dy << 16)
    int64_t pos;
    int count;
    int nrun;
    int run_end[kMaxRuns];
    __device__ inline void push(int c, int dx, int dy)
    {
        if (col) {
            col[pos] = c;
            sc[pos] = (int32_t)((uint32_t)(dx & 0xffff) | ((uint32_t)dy << 16));
            ++pos;
        }
        ++count;
    }
    __device__ inline void end_image()
    {
        if (nrun < kMaxRuns) run_end[nrun] = count;
        ++nrun;
    }
};

__device__ void search_intra(const GeomDev& g, int kh, PairSink& out)
{
    const int sh = g.i2g[4 * (size_t)kh + 3];
    const SheetDev& sd = g.s[sh];
    const int ih = g.i2g[4 * (size_t)kh + 0], jh = g.i2g[4 * (size_t)kh + 1], l0 = g.i2g[4 * (size_t)kh + 2];
    const double kp0 = g.lpos[3 * (size_t)kh], kp1 = g.lpos[3 * (size_t)kh + 1];
    const int S = sd.bc == 1 ? g.sc_search : 0;
    for (int dx = -S; dx < S + 1; ++dx)
        for (int dy = -S; dy < S + 1; ++dy) {
            double scd0 = 0.0, scd1 = 0.0;
            if (sd.bc == 1) {
                scd0 = dx * sd.sc[0][0] + dy * sd.sc[1][0];
                scd1 = dx * sd.sc[0][1] + dy * sd.sc[1][1];
            }
            const double p0 = kp0 + scd0, p1 = kp1 + scd1;
            const int i0 = clamp_cell(g.ainv[0][0] * p0 + g.ainv[0][1] * p1, sd.min0, sd.max0);
            const int j0 = clamp_cell(g.ainv[1][0] * p0 + g.ainv[1][1] * p1, sd.min1, sd.max1);
            const int W = sd.win;
            for (int i = i0 - W; i < i0 + W; ++i)
                for (int j = j0 - W; j < j0 + W; ++j)
                    for (int l = 0; l < g.norb; ++l) {
                        const int k2 = grid_at(sd, g.norb, i, j, l);
                        if (k2 == -1) continue;
                        int gd0 = i - ih, gd1 = j - jh;
                        if (sd.bc == 1) {
                            gd0 = gd0 - dx * sd.st[0][0] - dy * sd.st[1][0];
                            gd1 = gd1 - dx * sd.st[0][1] - dy * sd.st[1][1];
                        }
                        const double t = intralayer(sd.mat, l0, l, gd0, gd1);
                        if (t != 0 || (dx == 0 && dy == 0)) out.push(k2 + sd.base, dx, dy);
                    }
            out.end_image();
        }
}

__device__ void search_inter(const GeomDev& g, int kh, PairSink& out)
{
    const int sh = g.i2g[4 * (size_t)kh + 3];
    const double kp0 = g.i2p[3 * (size_t)kh], kp1 = g.i2p[3 * (size_t)kh + 1];
    const int S = g.bc == 1 ? g.sc_search : 0;
    const double cut2 = 12.5 * 12.5;
    const double zero = 0.0;
    for (int dx = -S; dx < S + 1; ++dx)
        for (int dy = -S; dy < S + 1; ++dy) {
            double scd0 = 0.0, scd1 = 0.0;
            if (g.bc == 1) {
                scd0 = dx * g.gsc[0][0] + dy * g.gsc[1][0];
                scd1 = dx * g.gsc[0][1] + dy * g.gsc[1][1];
            }
            const double p0 = kp0 + scd0, p1 = kp1 + scd1;
            for (int side = 0; side < 2; ++side) {
                int s0;
                if (side == 0) {
                    if (!(sh > 0)) continue;
                    s0 = sh - 1;
                } else {
                    if (!(sh < g.num_sheets - 1)) continue;
                    s0 = sh + 1;
                }
                const SheetDev& t = g.s[s0];
                // Hstruct::findNearest: rotate by -angle[s0] (cosines from the host libm)
                const double x_new = p0 * t.cneg - p1 * t.sneg + zero * g.lat[0][0] + zero * g.lat[1][0];
                const double y_new0 = p0 * t.sneg + p1 * t.cneg + zero * g.lat[0][1] + zero * g.lat[1][1];
                const double y_new1 = p1 * t.cneg + p0 * t.sneg + zero * g.lat[0][1] + zero * g.lat[1][1];
                const int i0 = clamp_cell(g.ainv[0][0] * x_new + g.ainv[0][1] * y_new0, t.min0, t.max0);
                const int j0 = clamp_cell(g.ainv[1][0] * x_new + g.ainv[1][1] * y_new1, t.min1, t.max1);
                const int ilo = max(0, i0 - g.inter_win), ihi = min(t.H, i0 + g.inter_win);
                const int jlo = max(0, j0 - g.inter_win), jhi = min(t.W, j0 + g.inter_win);
                for (int i = ilo; i < ihi; ++i)
                    for (int j = jlo; j < jhi; ++j)
                        for (int l = 0; l < g.norb; ++l) {
                            const int k2 = t.grid[((size_t)i * t.W + j) * g.norb + l];
                            if (k2 == -1) continue;
                            const size_t K = (size_t)k2 + t.base;
                            const double x2 = g.i2p[3 * K], y2 = g.i2p[3 * K + 1];
                            if ((x2 - p0) * (x2 - p0) + (y2 - p1) * (y2 - p1) < cut2) out.push((int)K, dx, dy);
                        }
            }
            out.end_image();
        }
}

__global__ __launch_bounds__(kThreads) void pair_count_kernel(GeomDev g, int32_t* cnt_intra, int64_t* cnt_total)
{
    const int kh = blockIdx.x * kThreads + threadIdx.x;
    if (kh >= g.n) return;
    PairSink a = {nullptr, nullptr, 0, 0}, b = {nullptr, nullptr, 0, 0};
    search_intra(g, kh, a);
    search_inter(g, kh, b);
    cnt_intra[kh] = a.count;
    cnt_total[kh] = (int64_t)a.count + b.count;
}

__device__ void sort_by_col(int32_t* col, int32_t* sc, int64_t n)
{
    // stable insertion sort == the reference's bubble sort permutation
    for (int64_t i = 1; i < n; ++i) {
        const int32_t c = col[i], s = sc[i];
        int64_t j = i - 1;
        while (j >= 0 && col[j] > c) {
            col[j + 1] = col[j];
            sc[j + 1] = sc[j];
            --j;
        }
        col[j + 1] = c;
        sc[j + 1] = s;
    }
}

__device__ inline int64_t distinct_runs(const int32_t* col, int64_t n)
{
    int64_t r = 0;
    for (int64_t i = 0; i < n; ++i)
        if (i == 0 || col[i] != col[i - 1]) ++r;
    return r;
}

// Stable merge of a list's per-image sorted runs (tcol/tsc) into col/sc: at equal
// columns the earlier image wins, which is the order the reference's stable
// sort by column leaves (Sheet::orderPairs / Hstruct::orderPairs).
__device__ void merge_runs(const int32_t* tcol, const int32_t* tsc, const PairSink& k, int32_t* col, int32_t* sc)
{
    int head[kMaxRuns];
    int start = 0;
    for (int r = 0; r < k.nrun; ++r) {
        head[r] = start;
        start = k.run_end[r];
    }
    for (int o = 0; o < k.count; ++o) {
        int best = -1;
        int32_t bc = 0;
        for (int r = 0; r < k.nrun; ++r)
            if (head[r] < k.run_end[r]) {
                const int32_t c = tcol[head[r]];
                if (best < 0 || c < bc) {
                    bc = c;
                    best = r;
                }
            }
        col[o] = bc;
        sc[o] = tsc[head[best]];
        ++head[best];
    }
}

// The pairs of one list, in the reference's order: merged from the image runs in
// the scratch copy when the list is sorted by column (periodic sheets), copied
// as found otherwise; more images than kMaxRuns fall back to the insertion sort.
__device__ void place_pairs(const int32_t* tcol, const int32_t* tsc, const PairSink& k, bool sorted, int32_t* col,
                            int32_t* sc)
{
    if (sorted && k.nrun <= kMaxRuns) {
        merge_runs(tcol, tsc, k, col, sc);
        return;
    }
    for (int o = 0; o < k.count; ++o) {
        col[o] = tcol[o];
        sc[o] = tsc[o];
    }
    if (sorted) sort_by_col(col, sc, k.count);
}

__global__ __launch_bounds__(kThreads) void pair_write_kernel(GeomDev g, const int64_t* pair_rp, const int32_t* cnt_intra,
                                                              int32_t* tcol, int32_t* tsc, int32_t* pcol, int32_t* psc,
                                                              int64_t* nslots)
{
    const int kh = blockIdx.x * kThreads + threadIdx.x;
    if (kh >= g.n) return;
    const int64_t base = pair_rp[kh];
    PairSink a = {tcol, tsc, base, 0};
    search_intra(g, kh, a);
    PairSink b = {tcol, tsc, base + a.count, 0};
    search_inter(g, kh, b);
    const int ni = cnt_intra[kh];
    const int64_t nt = pair_rp[kh + 1] - base;
    const int sh = g.i2g[4 * (size_t)kh + 3];
    place_pairs(tcol + base, tsc + base, a, g.s[sh].bc == 1, pcol + base, psc + base);
    place_pairs(tcol + base + ni, tsc + base + ni, b, g.bc == 1, pcol + base + ni, psc + base + ni);
    nslots[kh] = distinct_runs(pcol + base, ni) + distinct_runs(pcol + base + ni, nt - ni);
}

// ---- wavefront-per-row pair search with the lists staged in LDS -------------
// One wavefront per orbital row walks the same candidates as search_intra /
// search_inter, 64 at a time in the same order (image, i, j, l) and with the same
// arithmetic, so every keep decision is theirs.  Kept pairs are compacted in
// candidate order (ballot + lanes-below count) into the row's LDS lists, one run
// per image.  The write pass then places each pair straight at its position of
// the stable merge of the runs (periodic lists; earlier image first at equal
// columns): its index in its own run plus, for every other run, the number of
// entries that precede it there (<= its column for earlier runs, < for later
// ones), found by binary search in LDS.  The slot count of a row is the number of
// pairs whose column no earlier run holds.  No scratch copy of the pairs; rows
// with more pairs than the lists hold use the thread-per-row kernels above.
constexpr int kPairWaves = 4;                       // rows (wavefronts) per workgroup
constexpr int kPairCapIntra = 512;                  // staged intralayer pairs per row
constexpr int kPairCapInter = 1024;                 // staged interlayer pairs per row
constexpr int kPairCap = kPairCapIntra + kPairCapInter;

struct WaveLists {
    int32_t* col;       // [kPairCap]: intralayer list at 0, interlayer list at kPairCapIntra
    int32_t* sc;
    int32_t* run_end;   // [2][kMaxRuns + 1]: run ends of the intralayer, then the interlayer list
};

__device__ inline int lanes_below(uint64_t mask, int lane)
{
    return __popcll(mask & ((uint64_t(1) << lane) - 1));
}

// candidate q = (i n1 + j) n0 + l of a window of n1 columns and n0 orbitals, walked
// 64 at a time: lane k starts at q = k and advances by 64 per step with carries, so
// the per-step cost is a few adds and compares (the step's digits are divided out
// once per window)
struct WindowWalk {
    int l, j, i;       // this lane's digits
    int sl, sj, si;    // digits of 64
    int n0, n1;
    __device__ inline WindowWalk(int lane, int n0_, int n1_) : n0(n0_), n1(n1_)
    {
        l = lane % n0;
        j = (lane / n0) % n1;
        i = lane / (n0 * n1);
        sl = 64 % n0;
        sj = (64 / n0) % n1;
        si = 64 / (n0 * n1);
    }
    __device__ inline void next()
    {
        l += sl;
        const int c0 = l >= n0;
        l -= c0 ? n0 : 0;
        j += sj + c0;
        const int c1 = j >= n1;
        j -= c1 ? n1 : 0;
        i += si + c1;
    }
};

// kept candidates of one chunk appended to a list at cnt, and counted
__device__ inline void wave_push(bool keep, int col, int dx, int dy, int lane, int32_t* lc, int32_t* ls, int cap,
                                 int& cnt)
{
    const uint64_t m = __ballot(keep);
    if (keep) {
        const int pos = cnt + lanes_below(m, lane);
        if (pos < cap) {
            lc[pos] = col;
            ls[pos] = (int32_t)((uint32_t)(dx & 0xffff) | ((uint32_t)dy << 16));
        }
    }
    cnt += __popcll(m);
}

// the intralayer and interlayer candidates of row kh (search_intra, search_inter)
__device__ void wave_search(const GeomDev& g, int kh, int lane, const WaveLists& L, int& n_intra, int& n_inter,
                            int& runs_intra, int& runs_inter)
{
    const int sh = g.i2g[4 * (size_t)kh + 3];
    {
        const SheetDev& sd = g.s[sh];
        const int ih = g.i2g[4 * (size_t)kh + 0], jh = g.i2g[4 * (size_t)kh + 1], l0 = g.i2g[4 * (size_t)kh + 2];
        const double kp0 = g.lpos[3 * (size_t)kh], kp1 = g.lpos[3 * (size_t)kh + 1];
        const int S = sd.bc == 1 ? g.sc_search : 0;
        const int W = sd.win, norb = g.norb;
        const int per = 4 * W * W * norb;   // candidates of one image: (2W)^2 cells x orbitals
        int cnt = 0, nrun = 0;
        for (int dx = -S; dx < S + 1; ++dx)
            for (int dy = -S; dy < S + 1; ++dy) {
                double scd0 = 0.0, scd1 = 0.0;
                if (sd.bc == 1) {
                    scd0 = dx * sd.sc[0][0] + dy * sd.sc[1][0];
                    scd1 = dx * sd.sc[0][1] + dy * sd.sc[1][1];
                }
                const double p0 = kp0 + scd0, p1 = kp1 + scd1;
                const int i0 = clamp_cell(g.ainv[0][0] * p0 + g.ainv[0][1] * p1, sd.min0, sd.max0);
                const int j0 = clamp_cell(g.ainv[1][0] * p0 + g.ainv[1][1] * p1, sd.min1, sd.max1);
                WindowWalk ww(lane, norb, 2 * W);
                for (int b = 0; b < per; b += 64, ww.next()) {
                    const int q = b + lane;
                    bool keep = false;
                    int col = 0;
                    if (q < per) {
                        const int l = ww.l;
                        const int j = j0 - W + ww.j;
                        const int i = i0 - W + ww.i;
                        const int k2 = grid_at(sd, norb, i, j, l);
                        if (k2 != -1) {
                            int gd0 = i - ih, gd1 = j - jh;
                            if (sd.bc == 1) {
                                gd0 = gd0 - dx * sd.st[0][0] - dy * sd.st[1][0];
                                gd1 = gd1 - dx * sd.st[0][1] - dy * sd.st[1][1];
                            }
                            const double t = intralayer(sd.mat, l0, l, gd0, gd1);
                            keep = t != 0 || (dx == 0 && dy == 0);
                            col = k2 + sd.base;
                        }
                    }
                    wave_push(keep, col, dx, dy, lane, L.col, L.sc, kPairCapIntra, cnt);
                }
                if (lane == 0 && nrun < kMaxRuns) L.run_end[nrun] = cnt;
                ++nrun;
            }
        n_intra = cnt;
        runs_intra = nrun;
    }
    {
        const double kp0 = g.i2p[3 * (size_t)kh], kp1 = g.i2p[3 * (size_t)kh + 1];
        const int S = g.bc == 1 ? g.sc_search : 0;
        const double cut2 = 12.5 * 12.5;
        const double zero = 0.0;
        const int norb = g.norb;
        int cnt = 0, nrun = 0;
        for (int dx = -S; dx < S + 1; ++dx)
            for (int dy = -S; dy < S + 1; ++dy) {
                double scd0 = 0.0, scd1 = 0.0;
                if (g.bc == 1) {
                    scd0 = dx * g.gsc[0][0] + dy * g.gsc[1][0];
                    scd1 = dx * g.gsc[0][1] + dy * g.gsc[1][1];
                }
                const double p0 = kp0 + scd0, p1 = kp1 + scd1;
                for (int side = 0; side < 2; ++side) {
                    int s0;
                    if (side == 0) {
                        if (!(sh > 0)) continue;
                        s0 = sh - 1;
                    } else {
                        if (!(sh < g.num_sheets - 1)) continue;
                        s0 = sh + 1;
                    }
                    const SheetDev& t = g.s[s0];
                    const double x_new = p0 * t.cneg - p1 * t.sneg + zero * g.lat[0][0] + zero * g.lat[1][0];
                    const double y_new0 = p0 * t.sneg + p1 * t.cneg + zero * g.lat[0][1] + zero * g.lat[1][1];
                    const double y_new1 = p1 * t.cneg + p0 * t.sneg + zero * g.lat[0][1] + zero * g.lat[1][1];
                    const int i0 = clamp_cell(g.ainv[0][0] * x_new + g.ainv[0][1] * y_new0, t.min0, t.max0);
                    const int j0 = clamp_cell(g.ainv[1][0] * x_new + g.ainv[1][1] * y_new1, t.min1, t.max1);
                    const int ilo = max(0, i0 - g.inter_win), ihi = min(t.H, i0 + g.inter_win);
                    const int jlo = max(0, j0 - g.inter_win), jhi = min(t.W, j0 + g.inter_win);
                    const int nj = max(0, jhi - jlo);
                    const int per = max(0, ihi - ilo) * nj * norb;
                    WindowWalk ww(lane, norb, max(nj, 1));
                    for (int b = 0; b < per; b += 64, ww.next()) {
                        const int q = b + lane;
                        bool keep = false;
                        int col = 0;
                        if (q < per) {
                            const int l = ww.l;
                            const int j = jlo + ww.j;
                            const int i = ilo + ww.i;
                            const int k2 = t.grid[((size_t)i * t.W + j) * norb + l];
                            if (k2 != -1) {
                                const size_t K = (size_t)k2 + t.base;
                                const double x2 = g.i2p[3 * K], y2 = g.i2p[3 * K + 1];
                                keep = (x2 - p0) * (x2 - p0) + (y2 - p1) * (y2 - p1) < cut2;
                                col = (int)K;
                            }
                        }
                        wave_push(keep, col, dx, dy, lane, L.col + kPairCapIntra, L.sc + kPairCapIntra,
                                         kPairCapInter, cnt);
                    }
                }
                if (lane == 0 && nrun < kMaxRuns) L.run_end[kMaxRuns + 1 + nrun] = cnt;
                ++nrun;
            }
        n_inter = cnt;
        runs_inter = nrun;
    }
}

// entries of run [a, b) of lc below c (strict) or at most c
__device__ inline int run_rank(const int32_t* lc, int a, int b, int32_t c, bool or_equal)
{
    while (a < b) {
        const int m = (a + b) >> 1;
        if (lc[m] < c || (or_equal && lc[m] == c)) a = m + 1;
        else b = m;
    }
    return a;
}

// one list of a row placed at out (merged when sorted); returns its slot count
__device__ int wave_place(const int32_t* lc, const int32_t* ls, const int32_t* run_end, int cnt, int nrun, bool sorted,
                          int lane, int32_t* ocol, int32_t* osc)
{
    int slots = 0;
    for (int b = 0; b < cnt; b += 64) {
        const int e = b + lane;
        bool first = false;
        if (e < cnt) {
            const int32_t c = lc[e];
            if (!sorted) {
                ocol[e] = c;
                osc[e] = ls[e];
                first = e == 0 || lc[e - 1] != c;
            } else {
                // the run holding e
                int r = 0;
                while (run_end[r] <= e) ++r;
                int pos = e - (r == 0 ? 0 : run_end[r - 1]);
                int before_equal = 0;
                int start = 0;
                for (int q = 0; q < nrun; ++q) {
                    const int end = run_end[q];
                    if (q < r) {
                        const int le = run_rank(lc, start, end, c, true) - start;
                        const int lt = run_rank(lc, start, end, c, false) - start;
                        pos += le;
                        before_equal += le - lt;
                    } else if (q > r) {
                        pos += run_rank(lc, start, end, c, false) - start;
                    }
                    start = end;
                }
                ocol[pos] = c;
                osc[pos] = ls[e];
                first = before_equal == 0;
            }
        }
        slots += __popcll(__ballot(first));
    }
    return slots;
}

__global__ __launch_bounds__(64 * kPairWaves) void pair_write_wave_kernel(GeomDev g, const int64_t* pair_rp,
                                                                          int32_t* pcol, int32_t* psc, int64_t* nslots)
{
    __shared__ int32_t lcol[kPairWaves][kPairCap];
    __shared__ int32_t lsc[kPairWaves][kPairCap];
    __shared__ int32_t lrun[kPairWaves][2 * (kMaxRuns + 1)];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int kh = blockIdx.x * kPairWaves + w;
    if (kh >= g.n) return;
    WaveLists L = {lcol[w], lsc[w], lrun[w]};
    int ni = 0, nx = 0, ri = 0, rx = 0;
    wave_search(g, kh, lane, L, ni, nx, ri, rx);
    const int64_t base = pair_rp[kh];
    const int sh = g.i2g[4 * (size_t)kh + 3];
    // the host runs this kernel only when every list fits and has at most kMaxRuns runs
    if (ni > kPairCapIntra || nx > kPairCapInter || ri > kMaxRuns || rx > kMaxRuns || pair_rp[kh + 1] - base != ni + nx)
        return;
    const int si = wave_place(lcol[w], lsc[w], lrun[w], ni, ri, g.s[sh].bc == 1, lane, pcol + base, psc + base);
    const int sx = wave_place(lcol[w] + kPairCapIntra, lsc[w] + kPairCapIntra, lrun[w] + kMaxRuns + 1, nx, rx,
                              g.bc == 1, lane, pcol + base + ni, psc + base + ni);
    if (lane == 0) nslots[kh] = (int64_t)si + sx;
}

// ---- per-job fill -----------------------------------------------------------
__global__ __launch_bounds__(kThreads) void positions_kernel(GeomDev g, JobDev j, double* i2pos)
{
    // Locality::setConfigPositions (locality.cpp:3618-3643, 3849-3857)
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= g.n) return;
    const int s = g.i2g[4 * (size_t)i + 3];
    const double* a = j.s1a[s];   // rotated lattice of sheet s: a1 = (a[0], a[1]), a2 = (a[2], a[3])
    const double sh0 = j.shift[s][0], sh1 = j.shift[s][1];
    double x = g.i2p[3 * (size_t)i + 0] + sh0 * a[0] + sh1 * a[2];
    double y = g.i2p[3 * (size_t)i + 1] + sh0 * a[1] + sh1 * a[3];
    double z = g.i2p[3 * (size_t)i + 2];
    if (j.strain_type == 1 || j.strain_type == 6) {
        // relaxation field (locality.cpp:3657-3711; StrainCalc::fourierStrainDisp_sc and
        // fourierStrainDisp, strain.cpp:290-311, 390-460): directions k = p b1 + q b2,
        // q < max(p, 1), each with its three 60-degree turns (one for the constant term).
        // STRAIN_TYPE = DIRECT (6): the same sum over the built-in max_k = 2 coefficients of
        // StrainCalc::supercellDisp (strain.cpp:756-876)
        const bool direct = j.strain_type == 6;
        const int max_k = direct ? 2 : j.relax_max_k;
        const int nc = max_k * (max_k + 1) / 2 + 1;
        double ux = 0.0, uy = 0.0, uz = 0.0;
        int idx = 0;
        for (int p = 0; p < max_k + 1; ++p) {
            for (int q = 0; q < (p > 1 ? p : 1); ++q) {
                const double ff = p == 0 ? 1.0 : 2.0;
                const int rot_max = p == 0 ? 1 : 3;
                const double kb0 = p * j.rb[0][0] + q * j.rb[1][0], kb1 = p * j.rb[0][1] + q * j.rb[1][1];
                double cx, cy, cz;
                if (direct) {
                    cx = 0.0;
                    cy = idx == 1 ? -0.019951155162081 : idx == 2 ? -0.000806807799645 : 0.0;
                    cz = idx == 0 ? -0.061901691526695 : idx == 1 ? -0.0220634404118861
                       : idx == 2 ? -0.004100201789620 : -0.002676762110702;
                } else {
                    cx = j.relax[idx];
                    cy = j.relax[nc + idx];
                    cz = j.relax[2 * nc + idx];
                }
                for (int r = 0; r < rot_max; ++r) {
                    const double c = j.rot_cs[r][0], sn = j.rot_cs[r][1];
                    const double k0 = c * kb0 - sn * kb1, k1 = sn * kb0 + c * kb1;
                    const double c0 = c * cx - sn * cy, c1 = sn * cx + c * cy;
                    const double ph = k0 * x + k1 * y;
                    ux = ux + ff * c0 * sin(ph);
                    uy = uy + ff * c1 * sin(ph);
                    uz = uz + ff * cz * cos(ph);
                }
                ++idx;
            }
        }
        const double rx = j.basis_cs[0] * ux - j.basis_cs[1] * uy, ry = j.basis_cs[1] * ux + j.basis_cs[0] * uy;
        // per-sheet factors: SUPERCELL moves the second sheet the other way; DIRECT scales the
        // in-plane field by -2 on sheet 2 and the out-of-plane one by 0 on sheet 2 and -1 on
        // sheet 3 (factors of 0 and powers of two, so scaling the sum equals scaling the terms)
        const double fin = direct ? (s == 1 ? -2.0 : 1.0) : (s == 1 ? -1.0 : 1.0);
        const double fout = direct ? (s == 2 ? -1.0 : (s == 1 ? 0.0 : 1.0)) : (s == 1 ? -1.0 : 1.0);
        x = x + fin * rx;
        y = y + fin * ry;
        z = z + fout * uz;
    }
    if (j.sheet_strain_on == 1) {
        // per-sheet strain (open boundaries only), before the uniform strain
        const double tx = x, ty = y;
        x = (1.0 + j.ss[s][0][0]) * tx + j.ss[s][0][1] * ty;
        y = j.ss[s][1][0] * tx + (1.0 + j.ss[s][1][1]) * ty;
    }
    if (j.uniform_strain == 1) {
        const double tx = x, ty = y;
        x = (1.0 + j.u[0][0]) * tx + j.u[0][1] * ty;
        y = j.u[1][0] * tx + (1.0 + j.u[1][1]) * ty;
    }
    i2pos[3 * (size_t)i + 0] = x;
    i2pos[3 * (size_t)i + 1] = y;
    i2pos[3 * (size_t)i + 2] = z;
}

__global__ __launch_bounds__(kThreads) void local_theta_kernel(GeomDev g, const double* i2pos, double* theta)
{
    // Locality::getLocalTheta (locality.cpp:665-731): the first nearest neighbour
    // among the orbitals at index offsets -3, -1, 1, 3 (isNearestNeighbor,
    // locality.cpp:733-767: same sheet, the other orbital, the same cell or the cell
    // one step back (from A) / forward (from B) in i or j); the bond angle less
    // pi/6 (A) or pi/2 (B); the sheet's angle when none is found
    const int k1 = blockIdx.x * kThreads + threadIdx.x;
    if (k1 >= g.n) return;
    const int* a = g.i2g + 4 * (size_t)k1;
    const int step = a[2] == 0 ? -1 : 1;
    int k2 = -1;
    for (int dk = -3; dk < 4 && k2 < 0; dk += 2) {
        const int t = k1 + dk;
        if (t < 0 || t >= g.n) continue;
        const int* c = g.i2g + 4 * (size_t)t;
        if (c[3] != a[3] || c[2] == a[2]) continue;
        const int di = c[0] - a[0], dj = c[1] - a[1];
        if ((di == 0 && dj == 0) || (di == step && dj == 0) || (di == 0 && dj == step)) k2 = t;
    }
    double th = g.s[a[3]].angle;
    if (k2 >= 0) {
        const double dx = i2pos[3 * (size_t)k2] - i2pos[3 * (size_t)k1];
        const double dy = i2pos[3 * (size_t)k2 + 1] - i2pos[3 * (size_t)k1 + 1];
        th = atan2(dy, dx) - (a[2] == 0 ? 0.523598775598298873077 : 1.57079632679489661923);
    }
    theta[k1] = th;
}

template <bool CPX>
__global__ __launch_bounds__(kThreads) void fill_kernel(GeomDev g, JobDev j, const double* i2pos, const int64_t* pair_rp,
                                                        const int32_t* cnt_intra, const int32_t* pcol,
                                                        const int32_t* psc, const int64_t* slot_rp, int32_t* scol,
                                                        double* sval, uint8_t* skeep, double* sdx, double* sdy,
                                                        const double* ltheta)
{
    const int k_i = blockIdx.x * kThreads + threadIdx.x;
    if (k_i >= g.n) return;
    const int64_t q0 = pair_rp[k_i], q1 = pair_rp[k_i + 1];
    const int64_t qi = q0 + cnt_intra[k_i];
    int64_t slot = slot_rp[k_i];
    const int s_i = g.i2g[4 * (size_t)k_i + 3];
    const int mat = g.s[s_i].mat;
    const int ih = g.i2g[4 * (size_t)k_i + 0], jh = g.i2g[4 * (size_t)k_i + 1], l0 = g.i2g[4 * (size_t)k_i + 2];
    const double xi = i2pos[3 * (size_t)k_i], yi = i2pos[3 * (size_t)k_i + 1], zi = i2pos[3 * (size_t)k_i + 2];
    const double resc = j.energy_rescale, eshift = j.energy_shift;
    double ar = 0.0, ai = 0.0;
    // velocity slots (sdx != nullptr): (x1 - x2) t and (y1 - y2) t, real or complex like H
    double axr = 0.0, axi = 0.0, ayr = 0.0, ayi = 0.0;
    bool keep = false;
    for (int64_t q = q0; q < q1; ++q) {
        const bool intra = q < qi;
        const int new_k = pcol[q];
        const int32_t scp = psc[q];
        const int sc_i = (int)(int16_t)(scp & 0xffff), sc_j = (int)(int16_t)((uint32_t)scp >> 16);
        double psx = 0.0, psy = 0.0;
        if (g.bc == 1) {
            psx = sc_i * j.sc[0][0] + sc_j * j.sc[1][0];
            psy = sc_i * j.sc[0][1] + sc_j * j.sc[1][1];
        }
        const double x1 = xi + psx, y1 = yi + psy, z1 = zi;
        const double x2 = i2pos[3 * (size_t)new_k], y2 = i2pos[3 * (size_t)new_k + 1], z2 = i2pos[3 * (size_t)new_k + 2];
        double t;
        bool support;
        if (intra) {
            const SheetDev& sd = g.s[s_i];
            int gd0 = g.i2g[4 * (size_t)new_k + 0] - ih, gd1 = g.i2g[4 * (size_t)new_k + 1] - jh;
            if (sd.bc == 1) {
                gd0 = gd0 - sc_i * sd.st[0][0] - sc_j * sd.st[1][0];
                gd1 = gd1 - sc_i * sd.st[0][1] - sc_j * sd.st[1][1];
            }
            double raw_t = intralayer(mat, l0, g.i2g[4 * (size_t)new_k + 2], gd0, gd1);
            if (mat == MAT_STRAINED_GRAPHENE && (j.strain_type == 1 || j.strain_type == 6 || j.uniform_strain == 1)) {
                // strained intralayer term (locality.cpp:4034-4160, 4744-4870): the hop's direction on
                // the job's positions, the uniform strain (zero for the relaxation field), the sheet angle
                double hx = x2 - x1, hy = y2 - y1;
                const double hn = sqrt(hx * hx + hy * hy);
                if (hn != 0) {
                    hx = hx / hn;
                    hy = hy / hn;
                }
                const double zero_u[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
                raw_t = strained_intra(l0, g.i2g[4 * (size_t)new_k + 2], gd0, gd1, hx, hy,
                                       j.uniform_strain == 1 ? j.u : zero_u, sd.cneg, -sd.sneg);
            }
            if (mat == MAT_GRAPHENE_SIMPLE_STRAIN && raw_t != 0) {
                const double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
                raw_t = raw_t * exp(-kStrainBeta * (sqrt(dx * dx + dy * dy + dz * dz) / kAcc - 1.0));
            }
            const bool diag = CPX ? (new_k == k_i && sc_i == 0 && sc_j == 0) : (new_k == k_i);
            if (diag) {
                if (j.elec_on == 1) t = (raw_t + eshift + z1 * j.E) / resc;
                else t = (raw_t + eshift) / resc;
            } else {
                t = raw_t / resc;
            }
            support = raw_t != 0 || (diag && (eshift != 0 || j.elec_on == 1));
        } else {
            const int s2 = g.i2g[4 * (size_t)new_k + 3];
            const double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
            // with a relaxation field, each orbital's local twist (locality.cpp:4261-4269, 5045-5052)
            const double th1 = ltheta ? ltheta[k_i] : g.s[s_i].angle;
            const double th2 = ltheta ? ltheta[new_k] : g.s[s2].angle;
            t = interlayer(mat, l0, g.i2g[4 * (size_t)new_k + 2], dx, dy, dz, th1, th2) / resc;
            support = inter_support(mat, dx, dy, dz);
        }
        if (!CPX) {
            if (t != 0) {
                ar += t;
                axr += (x1 - x2) * t;   // locality.cpp:4182-4183, 4299-4300
                ayr += (y1 - y2) * t;
            }
        } else {
            // peierlsPhase (locality.cpp:7157-7161) and k.(r2 - r1) (locality.cpp:4896-4905)
            double phase = 3.14e-5 * j.B * (x2 - x1) * (0.5) * (y1 + y2);
            phase = phase + j.k[0] * (x2 - x1) + j.k[1] * (y2 - y1);
            double sn, cs;
            sincos(phase, &sn, &cs);
            const double tr = t * cs, ti = t * sn;
            if (tr != 0 || ti != 0) {
                ar += tr;
                ai += ti;
                axr += (x1 - x2) * tr;   // locality.cpp:4939-4940, 5113-5114
                axi += (x1 - x2) * ti;
                ayr += (y1 - y2) * tr;
                ayi += (y1 - y2) * ti;
            }
        }
        keep = keep || support;
        const bool last = (q + 1 == q1) || (q + 1 == qi) || pcol[q + 1] != new_k;
        if (last) {
            scol[slot] = new_k;
            if (CPX) {
                sval[2 * slot] = ar;
                sval[2 * slot + 1] = ai;
                if (sdx) {
                    sdx[2 * slot] = axr;
                    sdx[2 * slot + 1] = axi;
                    sdy[2 * slot] = ayr;
                    sdy[2 * slot + 1] = ayi;
                }
            } else {
                sval[slot] = ar;
                if (sdx) {
                    sdx[slot] = axr;
                    sdy[slot] = ayr;
                }
            }
            skeep[slot] = keep ? 1 : 0;
            ++slot;
            ar = 0.0;
            ai = 0.0;
            axr = axi = ayr = ayi = 0.0;
            keep = false;
        }
    }
}

__global__ __launch_bounds__(kThreads) void keep_count_kernel(int n, const int64_t* slot_rp, const uint8_t* skeep,
                                                              int64_t* cnt)
{
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    int64_t c = 0;
    for (int64_t q = slot_rp[r]; q < slot_rp[r + 1]; ++q) c += skeep[q];
    cnt[r] = c;
}

template <bool CPX>
__global__ __launch_bounds__(kThreads) void compact_kernel(int n, const int64_t* slot_rp, const uint8_t* skeep,
                                                           const int32_t* scol, const double* sval, const int64_t* out_rp,
                                                           int32_t* ocol, double* oval)
{
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    int64_t w = out_rp[r];
    for (int64_t q = slot_rp[r]; q < slot_rp[r + 1]; ++q) {
        if (!skeep[q]) continue;
        ocol[w] = scol[q];
        if (CPX) {
            oval[2 * w] = sval[2 * q];
            oval[2 * w + 1] = sval[2 * q + 1];
        } else {
            oval[w] = sval[q];
        }
        ++w;
    }
}

// exclusive scan of n int64 counts into out[0..n] (out[n] = total)
hipError_t scan_counts(const int64_t* cnt, int64_t* out, int n, hipStream_t s)
{
    size_t tmp_bytes = 0;
    hipError_t e = hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, cnt, out + 1, n, s);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    e = hipMallocAsync(&tmp, tmp_bytes + 16, s);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, cnt, out + 1, n, s);
    hipFreeAsync(tmp, s);
    if (e != hipSuccess) return e;
    return hipMemsetAsync(out, 0, sizeof(int64_t), s);
}

}  // namespace

hipError_t scan_exclusive_i64(const int64_t* cnt, int64_t* out, int n, hipStream_t s) { return scan_counts(cnt, out, n, s); }

hipError_t pairs_count(const GeomDev& g, int32_t* cnt_intra, int64_t* cnt_total, int64_t* pair_rp, hipStream_t s)
{
    hipLaunchKernelGGL(pair_count_kernel, rows_grid(g.n), dim3(kThreads), 0, s, g, cnt_intra, cnt_total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return scan_counts(cnt_total, pair_rp, g.n, s);
}

bool pairs_staged_fit(const GeomDev& g, int max_intra, int max_inter)
{
    return max_intra <= kPairCapIntra && max_inter <= kPairCapInter && g.sc_search >= 0 &&
           (2 * g.sc_search + 1) * (2 * g.sc_search + 1) <= kMaxRuns;
}

hipError_t pairs_write_staged(const GeomDev& g, const int64_t* pair_rp, int32_t* pcol, int32_t* psc, int64_t* nslots,
                              int64_t* slot_rp, hipStream_t s)
{
    hipLaunchKernelGGL(pair_write_wave_kernel, dim3((unsigned)((g.n + kPairWaves - 1) / kPairWaves)),
                       dim3(64 * kPairWaves), 0, s, g, pair_rp, pcol, psc, nslots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return scan_counts(nslots, slot_rp, g.n, s);
}

hipError_t pairs_write(const GeomDev& g, const int64_t* pair_rp, const int32_t* cnt_intra, int32_t* tcol, int32_t* tsc,
                       int32_t* pcol, int32_t* psc, int64_t* nslots, int64_t* slot_rp, hipStream_t s)
{
    hipLaunchKernelGGL(pair_write_kernel, rows_grid(g.n), dim3(kThreads), 0, s, g, pair_rp, cnt_intra, tcol, tsc, pcol,
                       psc, nslots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return scan_counts(nslots, slot_rp, g.n, s);
}

hipError_t job_positions(const GeomDev& g, const JobDev& j, double* i2pos, hipStream_t s)
{
    hipLaunchKernelGGL(positions_kernel, rows_grid(g.n), dim3(kThreads), 0, s, g, j, i2pos);
    return hipGetLastError();
}

hipError_t local_theta(const GeomDev& g, const double* i2pos, double* theta, hipStream_t s)
{
    hipLaunchKernelGGL(local_theta_kernel, rows_grid(g.n), dim3(kThreads), 0, s, g, i2pos, theta);
    return hipGetLastError();
}

hipError_t fill_slots(const GeomDev& g, const JobDev& j, const double* i2pos, const int64_t* pair_rp,
                      const int32_t* cnt_intra, const int32_t* pcol, const int32_t* psc, const int64_t* slot_rp,
                      int32_t* scol, double* sval, uint8_t* skeep, double* sdx, double* sdy, const double* ltheta,
                      hipStream_t s)
{
    if ((sdx == nullptr) != (sdy == nullptr)) return hipErrorInvalidValue;   // both velocity arrays or neither
    if (j.complex_matrix) {
        hipLaunchKernelGGL(fill_kernel<true>, rows_grid(g.n), dim3(kThreads), 0, s, g, j, i2pos, pair_rp, cnt_intra,
                           pcol, psc, slot_rp, scol, sval, skeep, sdx, sdy, ltheta);
    } else {
        hipLaunchKernelGGL(fill_kernel<false>, rows_grid(g.n), dim3(kThreads), 0, s, g, j, i2pos, pair_rp, cnt_intra,
                           pcol, psc, slot_rp, scol, sval, skeep, sdx, sdy, ltheta);
    }
    return hipGetLastError();
}

hipError_t prune_count(int n, const int64_t* slot_rp, const uint8_t* skeep, int64_t* cnt, int64_t* out_rp, hipStream_t s)
{
    hipLaunchKernelGGL(keep_count_kernel, rows_grid(n), dim3(kThreads), 0, s, n, slot_rp, skeep, cnt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return scan_counts(cnt, out_rp, n, s);
}

hipError_t prune_compact(int n, int is_complex, const int64_t* slot_rp, const uint8_t* skeep, const int32_t* scol,
                         const double* sval, const int64_t* out_rp, int32_t* ocol, double* oval, hipStream_t s)
{
    if (is_complex)
        hipLaunchKernelGGL(compact_kernel<true>, rows_grid(n), dim3(kThreads), 0, s, n, slot_rp, skeep, scol, sval,
                           out_rp, ocol, oval);
    else
        hipLaunchKernelGGL(compact_kernel<false>, rows_grid(n), dim3(kThreads), 0, s, n, slot_rp, skeep, scol, sval,
                           out_rp, ocol, oval);
    return hipGetLastError();
}

}  // namespace leo2d
