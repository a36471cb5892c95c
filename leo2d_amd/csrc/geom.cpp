// geom.cpp -- orbital enumeration, global index, positions, job list, writers.
//
// Enumeration follows Sheet::setIndex / posAtomGrid / checkShape
// (sheet.cpp:217-285, 482-559); the global index and rotated positions follow
// Hstruct::setIndex / posAtomIndex (hstruct.cpp:69-110, 150-218).  All of this
// is O(grid) host work done once per geometry; the O(N * window) pair search
// and the per-job fill run on the GPU (assemble.hip).
//
// Floating-point expressions keep the reference's operation order (and the
// library is built with -ffp-contract=off) so positions are bit-identical to
// the reference's on the same host, which makes the GPU pair pattern
// bit-identical too.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <sstream>

#include "host.h"

namespace leo2d {

namespace {

void pos_atom_grid(const Geom& g, int s, int gi, int gj, int l, double p[3])
{
    const SheetData& sd = g.sheets[s];
    const int i = gi + sd.min_shape[0];
    const int j = gj + sd.min_shape[1];
    double x = i * g.lat[0][0] + j * g.lat[1][0] + g.apos[l][0];
    double y = i * g.lat[0][1] + j * g.lat[1][1] + g.apos[l][1];
    const double z = g.apos[l][2];
    if (g.gs_on[s] == 1) {
        x = x + g.gshift[s][0] * g.lat[0][0] + g.gshift[s][1] * g.lat[1][0];
        y = y + g.gshift[s][0] * g.lat[0][1] + g.gshift[s][1] * g.lat[1][1];
    }
    p[0] = x;
    p[1] = y;
    p[2] = z;
}

bool check_shape(const Geom& g, int s, const double p[3])
{
    const SheetData& sd = g.sheets[s];
    if (sd.boundary_condition == 1) {
        const double delta = 0.0, grid_delta = 0.000001;
        const double* inv = &g.scinv[4 * s];
        const double i = (p[0] + delta) * inv[0] + (p[1] + delta) * inv[1];
        const double j = (p[0] + delta) * inv[2] + (p[1] + delta) * inv[3];
        return i >= 0.0 - grid_delta && i < 1.0 - grid_delta && j >= 0.0 - grid_delta && j < 1.0 - grid_delta;
    }
    return p[0] * p[0] + p[1] * p[1] < sd.max_R * sd.max_R;
}

}  // namespace

Geom build_geometry(const Deck& d)
{
    Geom g;
    const Params& o = d.opts;
    if (o.geti("solver_space") != 0) throw DeckError("momentum-space solver is not on this path");
    const long long st = o.geti("strain_type");
    if (st != 0 && st != 1 && st != 6)
        throw DeckError("STRAIN_TYPE " + std::to_string(st) + " is not on this path (NONE, SUPERCELL and DIRECT are)");
    if (st != 0 && o.geti("boundary_condition") != 1)
        throw DeckError("STRAIN_TYPE SUPERCELL and DIRECT need a periodic supercell");
    if (o.geti("mat_from_file") != 0) throw DeckError("MAT_FROM_FILE is not on this path");
    // keys whose observables or couplings lie outside this path: refused rather than ignored
    for (const char* k : {"chiral_on", "ballistic_transport", "mat_soc_on", "mlmc", "gsfe_z_on"})
        if (o.i.count(k) && o.geti(k) != 0) throw DeckError(std::string(k) + " = 1 is not on this path");
    g.num_sheets = (int)d.sheets.size();
    g.bc = (int)o.geti("boundary_condition");
    g.sc_search = (int)o.geti("sc_search_size");
    if (g.bc == 1) {
        auto it = o.dm.find("supercell");
        if (it == o.dm.end()) throw DeckError("periodic deck without a supercell");
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) g.gsc[a][b] = it->second[a][b];
    }
    // Materials::lattice / orbital_pos (graphene.h:78-85, strained_graphene.h:50-56); one lattice
    // for all sheets: StrainedGraphene on every sheet or on none
    const int mat0 = d.sheets.empty() ? MAT_GRAPHENE : d.sheets[0].mat;
    for (const SheetData& sd : d.sheets)
        if ((sd.mat == MAT_STRAINED_GRAPHENE) != (mat0 == MAT_STRAINED_GRAPHENE))
            throw DeckError("StrainedGraphene must be the material of every sheet or of none");
    const double A = 2.4768;
    g.lat[0][0] = 1. * A; g.lat[0][1] = 0.;
    g.lat[1][0] = .5 * A; g.lat[1][1] = lattice_s32(mat0) * A;
    g.apos[0][0] = 0.; g.apos[0][1] = 0.; g.apos[0][2] = 0.;
    g.apos[1][0] = 0.5 * A; g.apos[1][1] = orbital_s36(mat0) * A; g.apos[1][2] = 0;
    {
        const double a11 = g.lat[0][0], a12 = g.lat[0][1], a21 = g.lat[1][0], a22 = g.lat[1][1];
        const double det_a = a11 * a22 - a12 * a21;
        g.ainv[0][0] = a22 / det_a;
        g.ainv[0][1] = -a21 / det_a;
        g.ainv[1][0] = -a12 / det_a;
        g.ainv[1][1] = a11 / det_a;
    }
    g.sheets = d.sheets;
    g.angle = d.angles;
    g.height = d.heights;
    g.gs_on.assign(g.num_sheets, (int)o.geti("global_shifts_on"));
    g.gshift.assign(g.num_sheets, {0.0, 0.0});
    if (o.geti("global_shifts_on") == 1) {
        auto it = o.dm.find("global_shifts");
        if (it == o.dm.end() || (int)it->second.size() < g.num_sheets) throw DeckError("GLOBAL_SHIFTS missing");
        for (int s = 0; s < g.num_sheets; ++s) g.gshift[s] = {it->second[s][0], it->second[s][1]};
    }
    g.H.assign(g.num_sheets, 0);
    g.W.assign(g.num_sheets, 0);
    g.count.assign(g.num_sheets, 0);
    g.base.assign(g.num_sheets, 0);
    g.scinv.assign(4 * g.num_sheets, 0.0);
    g.grid.resize(g.num_sheets);
    long long total = 0;
    for (int s = 0; s < g.num_sheets; ++s) {
        const SheetData& sd = g.sheets[s];
        if (sd.boundary_condition == 1) {
            const double(*sc)[2] = sd.supercell;
            const double det = sc[0][0] * sc[1][1] - sc[0][1] * sc[1][0];
            g.scinv[4 * s + 0] = sc[1][1] / det;
            g.scinv[4 * s + 1] = -sc[1][0] / det;
            g.scinv[4 * s + 2] = -sc[0][1] / det;
            g.scinv[4 * s + 3] = sc[0][0] / det;
        }
        const int H = std::max(0, sd.max_shape[0] - sd.min_shape[0]);
        const int W = std::max(0, sd.max_shape[1] - sd.min_shape[1]);
        g.H[s] = H;
        g.W[s] = W;
        g.grid[s].assign((size_t)H * W * g.norb, -1);
        int k = 0;
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j)
                for (int l = 0; l < g.norb; ++l) {
                    double p[3];
                    pos_atom_grid(g, s, i, j, l, p);
                    if (check_shape(g, s, p)) g.grid[s][((size_t)i * W + j) * g.norb + l] = k++;
                }
        g.count[s] = k;
        g.base[s] = (int)total;
        total += k;
        if (total > 0x7fffffffLL) throw DeckError("more than 2^31 orbitals");
    }
    g.n = (int)total;
    g.i2g.assign(4 * (size_t)g.n, 0);
    g.i2p.assign(3 * (size_t)g.n, 0.0);
    g.lpos.assign(3 * (size_t)g.n, 0.0);
    for (int s = 0; s < g.num_sheets; ++s) {
        const double theta = g.angle[s];
        for (int i = 0; i < g.H[s]; ++i)
            for (int j = 0; j < g.W[s]; ++j)
                for (int l = 0; l < g.norb; ++l) {
                    const int k = g.grid[s][((size_t)i * g.W[s] + j) * g.norb + l];
                    if (k < 0) continue;
                    const size_t K = (size_t)k + g.base[s];
                    g.i2g[4 * K + 0] = i;
                    g.i2g[4 * K + 1] = j;
                    g.i2g[4 * K + 2] = l;
                    g.i2g[4 * K + 3] = s;
                    double p[3];
                    pos_atom_grid(g, s, i, j, l, p);
                    g.lpos[3 * K + 0] = p[0];
                    g.lpos[3 * K + 1] = p[1];
                    g.lpos[3 * K + 2] = p[2];
                    const double zero = 0.0;  // Hstruct shifts (unused, zero)
                    g.i2p[3 * K + 0] = p[0] * std::cos(theta) - p[1] * std::sin(theta) + zero * g.lat[0][0] + zero * g.lat[1][0];
                    g.i2p[3 * K + 1] = p[0] * std::sin(theta) + p[1] * std::cos(theta) + zero * g.lat[0][1] + zero * g.lat[1][1];
                    g.i2p[3 * K + 2] = p[2] + g.height[s];
                }
    }
    return g;
}

std::vector<int> target_list(const Deck& d, const Geom& g)
{
    // Hstruct::getTargetList, solver types 1 and 2 (hstruct.cpp:1056-1079)
    std::vector<int> out;
    const Params& o = d.opts;
    auto it = o.iv.find("target_sheets");
    std::vector<int> ts = it != o.iv.end() ? it->second : std::vector<int>{0};
    const int nts = (int)o.geti("num_target_sheets");
    for (int q = 0; q < nts && q < (int)ts.size(); ++q) {
        const int s = ts[q];
        if (s < 0 || s >= g.num_sheets) throw DeckError("TARGET_SHEETS out of range");
        const SheetData& sd = g.sheets[s];
        const int cx = (sd.max_shape[0] - sd.min_shape[0]) / 2;
        const int cy = (sd.max_shape[1] - sd.min_shape[1]) / 2;
        int c = -1;
        if (cx >= 0 && cx < g.H[s] && cy >= 0 && cy < g.W[s]) c = g.grid[s][((size_t)cx * g.W[s] + cy) * g.norb + 0];
        if (c == -1) throw DeckError("target sheet has no orbital in its centre cell");
        c += g.base[s];
        for (int orb = 0; orb < g.norb; ++orb) out.push_back(c + orb);
    }
    return out;
}

namespace {

std::vector<std::vector<double>> reciprocal(const std::vector<std::vector<double>>& a_in)
{
    // Locality::getReciprocal (locality.cpp:7215-7262)
    const double a[3][3] = {{a_in[0][0], a_in[0][1], 0.0}, {a_in[1][0], a_in[1][1], 0.0}, {0.0, 0.0, 1.0}};
    auto cross = [](const double* x, const double* y, int dim) {
        if (dim == 0) return x[1] * y[2] - x[2] * y[1];
        if (dim == 1) return x[2] * y[0] - x[0] * y[2];
        return x[0] * y[1] - x[1] * y[0];
    };
    double den1 = 0.0, den2 = 0.0;
    for (int j = 0; j < 3; ++j) {
        den1 += a[0][j] * cross(a[1], a[2], j);
        den2 += a[1][j] * cross(a[2], a[0], j);
    }
    std::vector<std::vector<double>> b(2, std::vector<double>(2, 0.0));
    for (int k = 0; k < 2; ++k) {
        b[0][k] = 2 * M_PI * cross(a[1], a[2], k) / den1;
        b[1][k] = 2 * M_PI * cross(a[2], a[0], k) / den2;
    }
    return b;
}

using KVec = std::array<double, 2>;

// Locality::getReciprocal(int s) (locality.cpp:7168-7213): sheet s's lattice turned
// by its angle (and strained by u_ij on a periodic cell with UNIFORM_STRAIN = 1)
std::vector<std::vector<double>> sheet_reciprocal(const Deck& d, int s)
{
    const Params& o = d.opts;
    const double A = 2.4768;
    const double uc[2][2] = {{1. * A, 0.}, {.5 * A, lattice_s32(d.sheets[s].mat) * A}};
    const double th = d.angles[s];
    std::vector<std::vector<double>> a(2, std::vector<double>(2));
    for (int i = 0; i < 2; ++i) {
        a[i][0] = uc[i][0] * std::cos(th) - uc[i][1] * std::sin(th);
        a[i][1] = uc[i][0] * std::sin(th) + uc[i][1] * std::cos(th);
    }
    if (o.geti("boundary_condition") == 1 && o.geti("uniform_strain") == 1) {
        const auto& u = o.dm.at("u_ij");
        for (auto& t : a) {
            const double x = t[0], y = t[1];
            t[0] = (1.0 + u[0][0]) * x + u[0][1] * y;
            t[1] = u[1][0] * x + (1.0 + u[1][1]) * y;
        }
    }
    return reciprocal(a);
}

// Piecewise-linear walk over the corners v[0] -> v[1] -> ... with parameter
// c = span i / nk: leg s covers s <= c <= s + 1 and the point is
// (s + 1 - c) v[s] + (c - s) v[s + 1]  (K_TYPE 1-3, locality.cpp:1760-1778,
// 1864-1884, 1946-1970)
std::vector<KVec> corner_walk(const std::vector<KVec>& v, double span, int nk)
{
    std::vector<KVec> pts;
    const int legs = (int)v.size() - 1;
    for (int i = 0; i < nk; ++i) {
        const double c = (span / (double)nk) * i;
        int s = legs - 1;
        for (int q = 0; q < legs - 1; ++q)
            if (c <= q + 1) { s = q; break; }
        const double hi = s + 1.0, lo = s;
        pts.push_back({(hi - c) * v[s][0] + (c - lo) * v[s + 1][0], (hi - c) * v[s][1] + (c - lo) * v[s + 1][1]});
    }
    return pts;
}

// The high-symmetry path of K_TYPE 1-5 for one base job (locality.cpp:1655-2249).
// Types 4 and 5 step along lines given by slope and intercept; their arithmetic
// keeps the reference's operation order so the k vectors agree bit for bit.
std::vector<KVec> k_path(const Deck& d)
{
    const Params& o = d.opts;
    const long long kt = o.geti("k_type");
    const int nk = (int)o.geti("num_k1");
    const int ns = (int)d.sheets.size();
    const KVec gam = {0.0, 0.0};
    if (kt == 1 || kt == 3) {
        // supercell zone: K = b2 turned by -30 deg over 2 cos 30, M = b2 / 2
        if (kt == 1 && ns < 2) throw DeckError("K_TYPE = 1 (twisted zone) needs two or more sheets");
        const auto b = reciprocal(o.dm.at("supercell"));
        const double f = 1.0 / (2.0 * std::cos(M_PI / 6));
        const KVec kp = {f * (std::cos(-M_PI / 6) * b[1][0] - std::sin(-M_PI / 6) * b[1][1]),
                         f * (std::sin(-M_PI / 6) * b[1][0] + std::cos(-M_PI / 6) * b[1][1])};
        const KVec mp = {b[1][0] / 2.0, b[1][1] / 2.0};
        if (kt == 1) return corner_walk({kp, gam, mp, kp}, 3.0, nk);
        const KVec kq = {std::cos(-M_PI / 3) * kp[0] - std::sin(-M_PI / 3) * kp[1],
                         std::sin(-M_PI / 3) * kp[0] + std::cos(-M_PI / 3) * kp[1]};
        return corner_walk({gam, mp, kp, kq, gam}, 4.0, nk);
    }
    if (kt == 2) {
        // zone of sheet 1: M = b1 / 2, K = M turned by 30 deg over cos 30
        const auto b = sheet_reciprocal(d, 0);
        const double h = (M_PI / 3.0) / 2.0;
        const KVec mp = {b[0][0] / 2.0, b[0][1] / 2.0};
        const KVec kp = {(mp[0] * std::cos(h) - mp[1] * std::sin(h)) / std::cos(h),
                         (mp[0] * std::sin(h) + mp[1] * std::cos(h)) / std::cos(h)};
        return corner_walk({gam, mp, kp, gam}, 3.0, nk);
    }
    std::vector<KVec> pts;
    if (kt == 4) {
        // (M, N) cell zone: Gamma -> K -> M -> Gamma, nk points per line
        if (o.geti("supercell_type") != 1) throw DeckError("K_TYPE = 4 needs SUPERCELL_M_N");
        const int M = (int)o.geti("m_supercell"), N = (int)o.geti("n_supercell");
        // two supercell groups (one or two sheets): the bilayer cell; three sheets: the
        // trilayer cell A = ((0, t), (-t, t)), t = M^2 + M N + N^2 (locality.cpp:2018-2040)
        if (ns > 3) throw DeckError("K_TYPE = 4 takes at most three sheets");
        const int tt = M * M + M * N + N * N;
        const int A[2][2] = {{ns < 3 ? (M > N ? N : M) : 0, ns < 3 ? (M > N ? M : N) : tt},
                             {ns < 3 ? (M > N ? -M : -N) : -tt, ns < 3 ? M + N : tt}};
        const double a = 2.4768;
        const double uc[2][2] = {{1. * a, 0.}, {.5 * a, lattice_s32(d.sheets[0].mat) * a}};
        std::vector<std::vector<double>> sc(2, std::vector<double>(2));
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) sc[r][c] = A[r][0] * uc[0][c] + A[r][1] * uc[1][c];
        const auto bs = reciprocal(sc);
        const double r3 = std::sqrt(3), al = M_PI / 6;
        const double rot[2][2] = {{std::cos(al), std::sin(al)}, {-std::sin(al), std::cos(al)}};
        const KVec mp = {0.5 * (bs[0][0] + bs[1][0]), 0.5 * (bs[0][1] + bs[1][1])};
        KVec kp = {0.0, 0.0};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) kp[i] = kp[i] + rot[i][j] * bs[1][j] / r3;
        const double s1 = (kp[1] - gam[1]) / (kp[0] - gam[0]);
        const double s2 = (mp[1] - kp[1]) / (mp[0] - kp[0]);
        const double s3 = (gam[1] - mp[1]) / (gam[0] - mp[0]);
        const double c1 = (kp[0] * gam[1] - gam[0] * kp[1]) / (kp[0] - gam[0]);
        const double c3 = (gam[0] * mp[1] - mp[0] * gam[1]) / (gam[0] - mp[0]);
        const double l1 = kp[0] - gam[0], l2 = mp[0] - kp[0], l3 = gam[0] - mp[0];
        for (int i = 0; i < 3 * nk; ++i) {
            if (i < nk) {
                pts.push_back({gam[0] + i * l1 / nk, gam[1] + s1 * i * l1 / nk + c1});
            } else if (i < 2 * nk) {
                const int t = i - nk;
                pts.push_back({kp[0] + t * l2 / nk, kp[1] + s2 * t * l2 / nk});
            } else {
                const int t = i - 2 * nk;
                if (mp[0] == mp[1]) pts.push_back({mp[0], mp[1] + t * (gam[1] - mp[1]) / nk});
                else pts.push_back({mp[0] + t * l3 / nk, mp[1] + s3 * t * l3 / nk + c3});
            }
        }
        return pts;
    }
    if (kt == 5) {
        // K points of sheets 1, 2, 3 (K_s = (2 b1 + b2) / 3 of each sheet), joined by lines
        if (ns < 3) throw DeckError("K_TYPE = 5 needs three sheets");
        KVec K[3];
        for (int s = 0; s < 3; ++s) {
            const auto b = sheet_reciprocal(d, s);
            for (int i = 0; i < 2; ++i) K[s][i] = 1.0 / 3.0 * (2.0 * b[0][i] + b[1][i]);
        }
        const double s1 = (K[1][1] - K[0][1]) / (K[1][0] - K[0][0]);
        const double s2 = (K[2][1] - K[1][1]) / (K[2][0] - K[1][0]);
        const double m1 = K[1][1] - s1 * K[1][0], m2 = K[2][1] - s2 * K[2][0];
        const double l1 = K[1][0] - K[0][0], l2 = K[2][0] - K[1][0];
        for (int i = 0; i < 2 * nk + 1; ++i) {
            if (i < nk) {
                const double x = K[0][0] + i * l1 / nk;
                pts.push_back({x, s1 * x + m1});
            } else if (i < 2 * nk) {
                const double x = K[1][0] + (i - nk) * l2 / nk;
                pts.push_back({x, s2 * x + m2});
            } else {
                pts.push_back(K[2]);
            }
        }
        return pts;
    }
    throw DeckError("K_TYPE " + std::to_string(kt) + " is not a k-point scheme");
}

}  // namespace

void reciprocal_vectors(const double a[2][2], double b[2][2])
{
    const auto r = reciprocal({{a[0][0], a[0][1]}, {a[1][0], a[1][1]}});
    for (int i = 0; i < 2; ++i)
        for (int k = 0; k < 2; ++k) b[i][k] = r[i][k];
}

std::vector<leo2d_job> build_jobs(const Deck& d, const Geom& g)
{
    const Params& o = d.opts;
    const int ns = g.num_sheets;
    const std::vector<int> targets = target_list(d, g);
    if ((int)targets.size() > LEO2D_MAX_TARGETS) throw DeckError("too many targets");
    leo2d_job proto = {};
    proto.energy_rescale = o.getd("energy_rescale");
    proto.energy_shift = o.getd("energy_shift");
    proto.elec_on = (int)o.geti("elecOn");
    proto.mag_on = (int)o.geti("magOn");
    proto.E = o.getd("E");
    proto.B = o.getd("B");
    proto.k_sampling = (int)o.geti("k_sampling");
    // locality.cpp:2906-2908
    proto.complex_matrix = (o.geti("magOn") == 1 || o.geti("solver_space") == 1 || o.geti("k_sampling") == 1 ||
                            o.geti("chiral_on") == 1 || o.geti("ballistic_transport") == 1 ||
                            o.geti("mat_soc_on") == 1) ? 1 : 0;
    proto.uniform_strain = (int)o.geti("uniform_strain");
    if (proto.uniform_strain == 1) {
        auto it = o.dm.find("u_ij");
        if (it == o.dm.end()) throw DeckError("UNIFORM_STRAIN = 1 needs UNIFORM_STRAIN_MAT");
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) proto.u[a][b] = it->second[a][b];
    }
    proto.strain_type = (int)o.geti("strain_type");
    proto.num_targets = (int)targets.size();
    for (size_t t = 0; t < targets.size(); ++t) proto.targets[t] = targets[t];
    if (auto hs = o.dm.find("heterostrain"); hs != o.dm.end()) {
        if (g.bc == 1) throw DeckError("HETEROSTRAIN needs open boundaries");
        const double eps = hs->second[0][0], nu = hs->second[0][1];
        const int top = ns - 1;
        proto.sheet_strain_on = 1;
        proto.sheet_strain[top][0][0] = eps;
        proto.sheet_strain[top][1][1] = -nu * eps;
    }

    std::vector<leo2d_job> jobs;
    const long long st = o.geti("solver_type");
    if (st == 1) {
        // recursiveShiftCalc (locality.cpp:1222-1269)
        const int nS = (int)o.geti("nShifts");
        const int nss = (int)o.geti("num_shift_sheets");
        auto it = o.iv.find("shift_sheets");
        if (nss > 0 && (it == o.iv.end() || (int)it->second.size() < nss)) throw DeckError("SHIFT_SHEETS missing");
        const std::vector<int> sss = nss > 0 ? it->second : std::vector<int>();
        for (int q = 0; q < nss; ++q)
            if (sss[q] < 0 || sss[q] >= ns) throw DeckError("SHIFT_SHEETS out of range");
        const int max_jobs = (int)std::pow(nS * nS, nss);
        std::vector<std::vector<double>> shifts(ns, std::vector<double>(3, 0.0));
        std::function<void(int)> rec = [&](int level) {
            if (level == 0) {
                leo2d_job j = proto;
                j.job_id = (int)jobs.size() + 1;
                j.max_jobs = max_jobs;
                for (int s = 0; s < ns; ++s)
                    for (int c = 0; c < 3; ++c) j.shift[s][c] = shifts[s][c];
                jobs.push_back(j);
                return;
            }
            const int tar = sss[level - 1];
            for (int i = 0; i < nS; ++i)
                for (int k = 0; k < nS; ++k) {
                    shifts[tar][0] = (1.0 / (double)(nS)) * i;
                    shifts[tar][1] = (1.0 / (double)(nS)) * k;
                    shifts[tar][2] = 0;
                    rec(level - 1);
                }
        };
        rec(nss);
    } else if (st == 6) {
        // FIXED_SHIFT: one job, sheet 1 unshifted and every other sheet at (1/3, 1/3)
        // (locality.cpp:1580-1613)
        leo2d_job j = proto;
        j.job_id = 1;
        j.max_jobs = 1;
        for (int s = 0; s < ns; ++s) {
            j.shift[s][0] = s == 0 ? 0.0 : 1.0 / 3.0;
            j.shift[s][1] = s == 0 ? 0.0 : 1.0 / 3.0;
            j.shift[s][2] = 0.0;
        }
        jobs.push_back(j);
    } else if (st == 2) {
        // linecut jobs (locality.cpp:1330-1388)
        const int max_jobs = (int)o.geti("nShifts");
        const int nlc = (int)o.geti("num_lc_points");
        const auto& lc = o.dm.at("lc_points");
        if ((int)lc.size() < nlc || nlc < 2) throw DeckError("LC_POINTS needs NUM_LC_POINTS >= 2 points");
        for (int i = 0; i < max_jobs; ++i) {
            const double x = (1.0 / ((double)max_jobs)) * i;
            const double total_x = x * (nlc - 1);
            const int idx = (int)std::floor(total_x);
            const double eff = std::fmod(total_x, 1);
            leo2d_job j = proto;
            j.job_id = i + 1;
            j.max_jobs = max_jobs;
            j.shift[ns - 1][0] = lc[idx][0] + eff * (lc[idx + 1][0] - lc[idx][0]);
            j.shift[ns - 1][1] = lc[idx][1] + eff * (lc[idx + 1][1] - lc[idx][1]);
            j.shift[ns - 1][2] = 0;
            jobs.push_back(j);
        }
    } else {
        throw DeckError("SOLVER_TYPE " + std::to_string(st) + " is not on this path (SQ, LC and FIXED_SHIFT are)");
    }
    if (o.geti("k_sampling") == 1 && o.geti("k_type") != 0) {
        if (g.bc != 1) throw DeckError("K_SAMPLING needs BOUNDARY_CONDITION = 1");
        const std::vector<KVec> pts = k_path(d);
        std::vector<leo2d_job> kj;
        for (const leo2d_job& base : jobs)
            for (const KVec& kv : pts) {
                leo2d_job j = base;
                j.job_id = (int)kj.size() + 1;
                j.k[0] = kv[0];
                j.k[1] = kv[1];
                kj.push_back(j);
            }
        jobs = kj;
    } else if (o.geti("k_sampling") == 1) {
        if (g.bc != 1) throw DeckError("K_SAMPLING needs BOUNDARY_CONDITION = 1");
        const int nk1 = (int)o.geti("num_k1"), nk2 = (int)o.geti("num_k2");
        const auto b = reciprocal(o.dm.at("supercell"));
        std::vector<leo2d_job> kj;
        for (const leo2d_job& base : jobs)
            for (int k1 = 0; k1 < nk1; ++k1)
                for (int k2 = 0; k2 < nk2; ++k2) {
                    leo2d_job j = base;
                    j.job_id = (int)kj.size() + 1;
                    j.k[0] = (k1 / (double)nk1) * b[0][0] + (k2 / (double)nk2) * b[1][0];
                    j.k[1] = (k1 / (double)nk1) * b[0][1] + (k2 / (double)nk2) * b[1][1];
                    j.k_grid[0] = k1 / (double)nk1;
                    j.k_grid[1] = k2 / (double)nk2;
                    kj.push_back(j);
                }
        jobs = kj;
    }
    return jobs;
}

// ---- writers ----------------------------------------------------------------

namespace {
std::string fx(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.6f", v);  // std::fixed, precision 6
    return b;
}
}  // namespace

std::string cheb_text(const Deck& d, const std::vector<leo2d_job>& jobs, const std::vector<std::vector<double>>& rho,
                      int M)
{
    // Param_tools::save + saveHeader, DoS branch (param_tools.cpp:19-80, 228-285, 467-500, 564-577)
    const Params& o = d.opts;
    const double a = o.getd("energy_rescale"), b = o.getd("energy_shift");
    std::ostringstream out;
    const int verbose = (int)o.geti("verbose_save");
    const int st = (int)o.geti("solver_type");
    const int ns = (int)o.geti("num_sheets");
    for (size_t q = 0; q < jobs.size(); ++q) {
        const leo2d_job& j = jobs[q];
        const std::vector<double>& r = rho[q];
        const int T = std::min(j.num_targets, (int)(r.size() / (size_t)M));
        if (verbose != 0) {
            out << "JOBID = " << j.job_id << " \n";
            if (st == 1 || st == 2) {
                out << "SHEET: SHIFT_X, SHIFT_Y, SHIFT_Z \n";
                for (int s = 0; s < ns; ++s)
                    out << s + 1 << "    : " << fx(j.shift[s][0]) << ", " << fx(j.shift[s][1]) << ", " << fx(j.shift[s][2])
                        << " \n";
                out << "NUM_TAR = " << T << "\n";
                if (T != 0) {
                    out << "TAR_LIST: ";
                    for (int t = 0; t < T - 1; ++t) out << j.targets[t] << ", ";
                    out << j.targets[T - 1] << " \n";
                } else {
                    out << "NO_TAR \n";
                }
            }
            if (o.geti("magOn") == 1 || o.geti("elecOn") == 1) {
                out << "MAG_ON  = " << o.geti("magOn") << ", B = " << fx(o.getd("B")) << " \n";
                out << "ELEC_ON = " << o.geti("elecOn") << ", E = " << fx(o.getd("E")) << " \n";
            }
        }
        if (o.geti("observable_type") == 1) {
            // conductivity (param_tools.cpp:287-317): VERBOSE_SAVE = 1 writes only the
            // header (its matrix block is commented out in the reference); 0 writes the
            // energies, a blank line, then per target M rows of M values and a blank line
            if (verbose == 0) {
                if (j.job_id == 1) {
                    for (int i = 0; i < M; ++i) out << fx(b + a * std::cos((i * 1.0 + 0.5) * M_PI / M)) << " ";
                    out << "\n";
                    out << "\n";
                }
                for (int t = 0; t < j.num_targets; ++t) {
                    for (int i = 0; i < M; ++i) {
                        for (int k = 0; k < M; ++k) out << fx(r[((size_t)t * M + i) * M + k]) << " ";
                        out << "\n";
                    }
                    out << "\n";
                }
            }
            continue;
        }
        if (verbose == 1) {
            if (j.job_id == 1) {
                out << "E: ";
                for (int i = 0; i < M; ++i) out << fx(b + a * std::cos((i * 1.0 + 0.5) * M_PI / M)) << " ";
                out << "\n";
            }
            out << "T: \n";
            for (int t = 0; t < T; ++t) {
                out << j.targets[t] << ": ";
                for (int i = 0; i < M - 1; ++i) out << fx(r[(size_t)t * M + i]) << ", ";
                out << fx(r[(size_t)t * M + M - 1]) << "\n";
            }
            out << "\n";
        } else {
            if (j.job_id == 1) {
                for (int i = 0; i < M; ++i) out << fx(b + a * std::cos((i * 1.0 + 0.5) * M_PI / M)) << " ";
                out << "\n";
            }
            const int rows = o.geti("kpm_trace") == 1 ? 1 : T;
            for (int t = 0; t < rows; ++t) {
                for (int i = 0; i < M - 1; ++i) out << fx(r[(size_t)t * M + i]) << " ";
                out << fx(r[(size_t)t * M + M - 1]) << "\n";
            }
        }
    }
    return out.str();
}

std::string matrix_dat_text(int n, const int64_t* rp, const int32_t* col, const double* val, int is_complex)
{
    // locality.cpp:4366-4383 (real: |v| > 1e-12 only), 5213-5228 (complex: every slot);
    // default ostream float format == printf %g
    std::string out;
    char b[128];
    for (int i = 0; i < n; ++i)
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
            if (is_complex) {
                std::snprintf(b, sizeof b, "%d, %d, %g, %g\n", col[q] + 1, i + 1, val[2 * q], val[2 * q + 1]);
                out += b;
            } else if (std::fabs(val[q]) > 1e-12) {
                std::snprintf(b, sizeof b, "%d, %d, %g\n", col[q] + 1, i + 1, val[q]);
                out += b;
            }
        }
    return out;
}

std::string kpts_text(const std::vector<leo2d_job>& jobs)
{
    // Param_tools::saveKpts (param_tools.cpp:579-602): fixed with 10 decimals, a space
    // before a component that is >= 0, two tabs between the components, " \n" after
    std::string out;
    char b[128];
    for (const leo2d_job& j : jobs) {
        if (j.k_sampling != 1) continue;
        std::snprintf(b, sizeof b, "%s%.10f\t\t%s%.10f \n", j.k[0] >= 0 ? " " : "", j.k[0], j.k[1] >= 0 ? " " : "",
                      j.k[1]);
        out += b;
    }
    return out;
}

std::string pos_dat_text(int n, const double* i2pos, const int32_t* i2g)
{
    // _pos.dat (locality.cpp:4425-4443): std::fixed, precision 12, "x    y    z    sheet    orbital"
    std::string out;
    char b[256];
    for (int i = 0; i < n; ++i) {
        std::snprintf(b, sizeof b, "%.12f    %.12f    %.12f    %d    %d\n", i2pos[3 * (size_t)i], i2pos[3 * (size_t)i + 1],
                      i2pos[3 * (size_t)i + 2], i2g[4 * (size_t)i + 3], i2g[4 * (size_t)i + 2]);
        out += b;
    }
    return out;
}

std::string pair_pos_text(int64_t np, const int32_t* ij, const double* i2pos)
{
    // _intra_pos.dat / _inter_pos.dat (locality.cpp:4446-4475): default stream format (%g),
    // "x1, y1, z1, x2, y2, z2, i, j" for every pair of the list
    std::string out;
    char b[320];
    for (int64_t q = 0; q < np; ++q) {
        const double* p1 = i2pos + 3 * (size_t)ij[2 * q];
        const double* p2 = i2pos + 3 * (size_t)ij[2 * q + 1];
        std::snprintf(b, sizeof b, "%g, %g, %g, %g, %g, %g, %d, %d\n", p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], ij[2 * q],
                      ij[2 * q + 1]);
        out += b;
    }
    return out;
}

}  // namespace leo2d
