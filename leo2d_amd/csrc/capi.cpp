// capi.cpp -- implementation of the C-ABI in include/leo2d.h.
//
// One leo2d_handle = one HIP stream + the device state of one worker: the
// uploaded geometry, the pair lists (built once, like constructGeom), the
// current Hamiltonian (rebuilt per job, like workerChebSolve) and scratch for
// the KPM recursion.  No exception crosses the ABI: every entry point catches
// and converts to a LEO2D_ERR_* code with a message in leo2d_last_error().
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <thread>
#include <vector>


#include "assemble.h"
#include "devbuf.h"
#include "host.h"
#include "kernels.h"

struct leo2d_deck {
    leo2d::Deck d;
};

struct leo2d_geom {
    leo2d::Deck deck;   // after supercell setup
    leo2d::Geom g;
};

namespace {

using leo2d::DevBuf;
using leo2d::HostBuf;
using leo2d::HipError;
using leo2d::check;
using leo2d::dalloc;

struct ArgError : std::runtime_error {
    explicit ArgError(const std::string& m) : std::runtime_error(m) {}
};
struct StateError : std::runtime_error {
    explicit StateError(const std::string& m) : std::runtime_error(m) {}
};
struct NanError : std::runtime_error {
    explicit NanError(const std::string& m) : std::runtime_error(m) {}
};

// the SELL-4-sigma step with LDS-staged row tiles (sell.hip)
constexpr int kVariantSell = 20;

// Below this many rows the SELL tiles (at most 56 rows each) are fewer than the
// GPU's CUs, and a matrix assembled for one short run pays the host-side tile
// build for little: the vector-chunk kernel runs those.
constexpr int kSellMinRows = 16384;

// step kernel for a block of B vectors on n rows (LEO2D_KPM_VARIANT = 0, 1, 14,
// 15 or 20 overrides the default).  In the spatial order: B = 16, 32 or 64 run
// the SELL-4-sigma step with LDS-staged row tiles (20) where its layout applies
// (complex vectors) and n >= kSellMinRows, else B = 16 or 32 vector chunks split
// over the XCDs (14).  In the natural order (or where neither applies): B >= 64
// the wavefront-per-row kernel (1), or for the FP32 recursion at B = 64 two rows
// per wavefront with 16-byte gathers (15); other B lane per vector (0).
int kpm_variant(bool mixed, bool spatial, int B, int n)
{
    if (const char* e = std::getenv("LEO2D_KPM_VARIANT")) {
        const int v = std::atoi(e);
        if (v == 0 || v == 1 || v == 14 || v == 15 || v == kVariantSell) return v;
    }
    if (spatial && (B == 16 || B == 32)) return n >= kSellMinRows ? kVariantSell : 14;
    if (spatial && B >= 64) return kVariantSell;
    return mixed && spatial ? 15 : 1;
}

int env_int(const char* k, int d)
{
    const char* e = std::getenv(k);
    return e ? std::atoi(e) : d;
}

// row schedule of the wavefront-per-row kernels: 1 = one contiguous slice of
// rows per XCD (LEO2D_KPM_SCHED overrides the default)
int kpm_sched()
{
    const char* e = std::getenv("LEO2D_KPM_SCHED");
    return e ? std::atoi(e) : 0;
}

int next_pow2(int x)
{
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// SELL-4-sigma row tiles of a matrix for the LDS-staged step (sell.hip)
struct SellTiles {
    bool valid = false, failed = false, spatial = false;
    int ucap = 0, tmax = 0, ntiles = 0, ngroups = 0, tile_rows = 0, max_len = 0;
    int64_t entries = 0;
    double build_ms = 0.0;
    DevBuf row0, u0, ucols, g0, grow, glen, goff, eu, map, val, val32;
    bool val_valid = false, val32_valid = false;
    std::vector<int32_t> job_tile0;   // first tile of each job when cut at job boundaries
    DevBuf djob_tile0;
};

}  // namespace

struct leo2d_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // geometry
    bool has_geom = false;
    leo2d::Geom geom;
    leo2d::GeomDev gdev{};
    std::vector<DevBuf> grids;
    DevBuf i2g, i2p, lpos;
    // pairs
    bool has_pairs = false;
    bool pairs_staged = false;   // the last pair lists came from the LDS-staged wavefront kernels
    int64_t n_pairs = 0, n_slots = 0, n_intra = 0;
    DevBuf cnt_intra, cnt_tmp, pair_rp, pcol, psc, nslots, slot_rp;
    // matrix
    bool has_csr = false;
    int n = 0;
    int64_t nnz = 0;
    int is_complex = 0;
    DevBuf rp, col, val, val32, i2pos;
    bool val32_valid = false;
    // x-velocity matrix dxH on H's pattern and row order (leo2d_assemble_velocity)
    DevBuf val_dx, val_dy;     // velocity matrices on H's pattern, real or complex like H
    bool has_dx = false;
    // Fourier relaxation field (leo2d_set_relaxation) and the per-orbital local twist
    DevBuf relax, ltheta;
    int relax_max_k = -1;
    struct CondTiming {
        double recursion_ms = 0.0, gemm_ms = 0.0;
        int32_t block = 0;
    } cond_timing;
    // spatial order used inside the recursion (perm[orig] = new row)
    bool csr_geom_order = false;       // device H rows are geometry orbitals
    std::vector<int32_t> perm_host;
    DevBuf perm, iperm;
    bool pcsr_valid = false, pval32_valid = false;
    DevBuf prp, pcol_k, pval, pval32;
    // SELL-4-sigma row tiles of the recursion matrix for the LDS-staged step
    // (sell.hip): pattern built on the host on first use, values gathered on the device
    SellTiles sell;
    HostBuf sell_stage;   // pinned staging of a pattern for its SELL tile build, kept for the next build
    leo2d_timing timing{};
    std::vector<hipEvent_t> events;
    // batched jobs: stacked block-diagonal CSR of every appended Hamiltonian
    struct Batch {
        struct Pending {
            DevBuf rp, col, val;
            int n = 0;
            int64_t nnz = 0;
        };
        std::vector<Pending> pending;          // appended, not yet stacked
        int is_complex = -1;
        int64_t rows = 0, nnz = 0;             // stacked so far
        std::vector<int64_t> job_row0{0};      // [jobs + 1]
        DevBuf rp, col, val, val32, drow0;
        bool val32_valid = false;
        int tile_rows = 0, ntiles = 0;          // current tiling
        std::vector<int32_t> tile_row0, job_tile0;
        SellTiles sell;                         // SELL tiles of the stack, cut at job boundaries
        DevBuf dtile_row0, djob_tile0;
        // autotuned launch shape, per (R, precision, vector element type) of the current stack
        struct Tuned {
            int R, precision, vec_type, mode, tile_rows, grid;
        };
        std::vector<Tuned> tuned;
        // captured recursion (valid while the key matches)
        hipGraphExec_t exec = nullptr;
        struct GraphKey {
            int M, R, precision, mode, tile_rows, grid;
            uint64_t seed;
            const void *A, *B, *part, *mu, *val, *targets;
            int copy_wave;
            bool operator==(const GraphKey& o) const
            {
                return M == o.M && R == o.R && precision == o.precision && mode == o.mode && tile_rows == o.tile_rows &&
                       grid == o.grid && seed == o.seed && A == o.A && B == o.B && part == o.part && mu == o.mu &&
                       val == o.val && targets == o.targets && copy_wave == o.copy_wave;
            }
        } key{};
        DevBuf A, B, part, mu, tgt;
        leo2d_batch_timing timing{};
        int jobs() const { return (int)job_row0.size() - 1 + (int)pending.size(); }
        void drop_graph()
        {
            if (exec) (void)hipGraphExecDestroy(exec);
            exec = nullptr;
        }
    } batch;

    ~leo2d_handle()
    {
        for (auto& b : grids) b.release();
        for (DevBuf* b : {&i2g, &i2p, &lpos, &cnt_intra, &cnt_tmp, &pair_rp, &pcol, &psc, &nslots, &slot_rp, &rp, &col,
                          &val, &val32, &i2pos, &perm, &iperm, &prp, &pcol_k, &pval, &pval32})
            b->release();
        batch.drop_graph();
        for (auto e : events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename F>
int guarded(leo2d_handle* h, F&& f)
{
    try {
        if (h) {
            check(hipSetDevice(h->device), "hipSetDevice");
            // launch checks read hipGetLastError: start from a clean slate
            (void)hipGetLastError();
        }
        return f();
    } catch (const HipError& e) {
        if (h) h->err = e.what();
        return e.code;
    } catch (const leo2d::DeckError& e) {
        if (h) h->err = e.what();
        return LEO2D_ERR_DECK;
    } catch (const ArgError& e) {
        if (h) h->err = e.what();
        return LEO2D_ERR_ARG;
    } catch (const StateError& e) {
        if (h) h->err = e.what();
        return LEO2D_ERR_STATE;
    } catch (const NanError& e) {
        if (h) h->err = e.what();
        return LEO2D_ERR_NAN;
    } catch (const std::bad_alloc&) {
        if (h) h->err = "host allocation failed";
        return LEO2D_ERR_ALLOC;
    } catch (const std::exception& e) {
        if (h) h->err = e.what();
        return LEO2D_ERR_ARG;
    }
}

hipEvent_t event_at(leo2d_handle* h, size_t i)
{
    while (h->events.size() <= i) {
        hipEvent_t e;
        check(hipEventCreate(&e), "hipEventCreate");
        h->events.push_back(e);
    }
    return h->events[i];
}

void to_device(DevBuf& b, const void* src, size_t bytes, hipStream_t s, const char* what)
{
    dalloc(b, bytes, what);
    if (bytes) check(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s), what);
}

leo2d::JobDev job_dev(const leo2d::Geom& g, const leo2d_job& j)
{
    leo2d::JobDev d{};
    for (int s = 0; s < g.num_sheets; ++s) {
        d.shift[s][0] = j.shift[s][0];
        d.shift[s][1] = j.shift[s][1];
        // rotated lattice of the sheet (locality.cpp:3633-3638), host libm
        const double th = g.angle[s];
        for (int dd = 0; dd < 2; ++dd) {
            d.s1a[s][2 * dd + 0] = std::cos(th) * g.lat[dd][0] - std::sin(th) * g.lat[dd][1];
            d.s1a[s][2 * dd + 1] = std::sin(th) * g.lat[dd][0] + std::cos(th) * g.lat[dd][1];
        }
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) d.sc[a][b] = g.gsc[a][b];
    if (g.bc == 1 && j.uniform_strain == 1) {
        // locality.cpp:2947-2960
        for (int i = 0; i < 2; ++i) {
            d.sc[i][0] = (1.0 + j.u[0][0]) * g.gsc[i][0] + j.u[0][1] * g.gsc[i][1];
            d.sc[i][1] = j.u[1][0] * g.gsc[i][0] + (1.0 + j.u[1][1]) * g.gsc[i][1];
        }
    }
    d.energy_rescale = j.energy_rescale;
    d.energy_shift = j.energy_shift;
    d.elec_on = j.elec_on;
    d.E = j.E;
    d.B = j.B;
    d.k[0] = j.k[0];
    d.k[1] = j.k[1];
    d.complex_matrix = j.complex_matrix;
    d.uniform_strain = j.uniform_strain;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) d.u[a][b] = j.u[a][b];
    d.sheet_strain_on = j.sheet_strain_on;
    for (int s = 0; s < LEO2D_MAX_SHEETS; ++s)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) d.ss[s][a][b] = j.sheet_strain[s][a][b];
    d.strain_type = j.strain_type;
    if (j.strain_type == 1 || j.strain_type == 6) {
        // the relaxation field is periodic on the supercell (strain.cpp:390-460): its
        // directions are the supercell's reciprocal vectors turned by r pi / 3
        leo2d::reciprocal_vectors(g.gsc, d.rb);
        for (int r = 0; r < 3; ++r) {
            const double th = r * 3.14159265358979323846 / 3.0;
            d.rot_cs[r][0] = std::cos(th);
            d.rot_cs[r][1] = std::sin(th);
        }
        const double rot = std::atan2(d.rb[0][1], d.rb[0][0]);
        d.basis_cs[0] = std::cos(rot);
        d.basis_cs[1] = std::sin(rot);
    }
    return d;
}

void set_csr_owned(leo2d_handle* h, int n, int64_t nnz, int cpx)
{
    h->n = n;
    h->nnz = nnz;
    h->is_complex = cpx;
    h->has_csr = true;
    h->val32_valid = false;
    h->pcsr_valid = false;
    h->pval32_valid = false;
    h->sell.valid = false;
    h->sell.failed = false;
    h->sell.val_valid = false;
    h->sell.val32_valid = false;
    h->has_dx = false;
}

// Tiles of the rows [lo, hi), with tile starts forced at lo and at the rows
// marked in brk.  With `local`, every column lies in [lo, hi) (checked) and the
// per-column scratch covers only those rows; otherwise it covers all n.  Offsets
// (u0, g0, goff) are relative to this part's own arrays; rows, columns and CSR
// positions are global.
struct SellPart {
    std::vector<int32_t> row0, ucols, g0, grow, glen;
    std::vector<int64_t> u0, goff, map;
    std::vector<uint16_t> eu;
    int tile_rows = 0, max_len = 0;
    bool ok = true;
};

void build_sell_part(SellPart& P, const int64_t* rp, const int32_t* col, const uint8_t* brk, int lo, int hi, int ucap,
                     int tmax, bool local, int n)
{
    const int base = local ? lo : 0;
    const int m = local ? hi - lo : n;
    std::vector<int32_t> stamp(m, -1), loc(m, 0), added, halo, order;
    int r = lo, tile = 0;
    while (r < hi) {
        const int start = r;
        int size = 0;   // distinct columns of the tile so far (own rows included)
        while (r < hi && r - start < tmax && !(r > start && brk[r])) {
            if (rp[r + 1] - rp[r] > leo2d::kSellMaxLen) {
                P.ok = false;
                return;
            }
            added.clear();
            auto take = [&](int c) {
                if (stamp[c - base] != tile) {
                    stamp[c - base] = tile;
                    added.push_back(c);
                }
            };
            take(r);
            for (int64_t q = rp[r]; q < rp[r + 1]; ++q) {
                if (local && (col[q] < lo || col[q] >= hi)) {
                    P.ok = false;
                    return;
                }
                take(col[q]);
            }
            if (size + (int)added.size() > ucap) {
                for (int c : added) stamp[c - base] = -1;
                break;
            }
            size += (int)added.size();
            ++r;
        }
        if (r == start) {
            P.ok = false;
            return;
        }
        const int T = r - start;
        P.tile_rows = std::max(P.tile_rows, T);
        // union: own rows, then the other columns in increasing order
        halo.clear();
        for (int rr = start; rr < r; ++rr)
            for (int64_t q = rp[rr]; q < rp[rr + 1]; ++q) {
                const int c = col[q];
                if ((c < start || c >= r) && stamp[c - base] == tile) {
                    stamp[c - base] = -2 - tile;   // listed once
                    halo.push_back(c);
                }
            }
        std::sort(halo.begin(), halo.end());
        P.row0.push_back(start);
        P.u0.push_back((int64_t)P.ucols.size());
        for (int rr = start; rr < r; ++rr) {
            loc[rr - base] = rr - start;
            P.ucols.push_back(rr);
        }
        for (size_t k = 0; k < halo.size(); ++k) {
            loc[halo[k] - base] = T + (int)k;
            P.ucols.push_back(halo[k]);
        }
        // SELL-4 groups, rows by decreasing length (sigma = the tile)
        order.resize(T);
        for (int k = 0; k < T; ++k) order[k] = start + k;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
        P.g0.push_back((int32_t)P.glen.size());
        for (int k = 0; k < T; k += 4) {
            int L = 0;
            for (int j = 0; j < 4; ++j) {
                const int row = k + j < T ? order[k + j] : -1;
                P.grow.push_back(row);
                if (row >= 0) L = std::max(L, (int)(rp[row + 1] - rp[row]));
            }
            L = (L + 3) / 4 * 4;
            P.max_len = std::max(P.max_len, L);
            P.glen.push_back(L);
            P.goff.push_back((int64_t)P.eu.size());
            for (int e = 0; e < L; ++e)
                for (int j = 0; j < 4; ++j) {
                    const int row = k + j < T ? order[k + j] : -1;
                    if (row >= 0 && e < rp[row + 1] - rp[row]) {
                        const int64_t q = rp[row] + e;
                        P.eu.push_back((uint16_t)loc[col[q] - base]);
                        P.map.push_back(q);
                    } else {
                        P.eu.push_back(0);
                        P.map.push_back(-1);
                    }
                }
        }
        ++tile;
    }
}

// SELL-4-sigma row tiles for the LDS-staged step (sell.hip).  Greedy tiles of
// consecutive rows of the recursion matrix (at most kSellMaxRows) while the
// tile's distinct columns -- its own rows first, then the other columns in
// increasing order -- fit `ucap` LDS rows.  Within a tile the rows are sorted by
// length (longest first, ties by row) and cut into groups of four; a group of
// longest row L stores round_up(L, 4) entries per slot, entry e of slot s at
// goff + 4 e + s, padded with (union index 0, value 0).  `map` gives each
// stored entry's position in the CSR (-1 for padding) so the values can be
// gathered on the device for every matrix with this pattern.  Returns false
// when a row has more than kSellMaxLen entries or does not fit `ucap` alone.
// With `breaks` (ascending first rows of stacked jobs, ending with n) no tile
// crosses a job boundary, and job_tile0 lists each job's first tile.
bool build_sell_tiles(SellTiles& S, HostBuf& stage, hipStream_t s, int n, int64_t nnz, const int64_t* drp,
                      const int32_t* dcol, int ucap, int tmax, bool spatial, const std::vector<int64_t>* breaks)
{
    if (S.valid && S.ucap == ucap && S.tmax == tmax && S.spatial == spatial) return true;
    if (S.failed && S.ucap == ucap && S.tmax == tmax && S.spatial == spatial) return false;
    const auto t0 = std::chrono::steady_clock::now();
    // the pattern comes over into pinned memory the handle keeps: batched jobs
    // rebuild their tiles for every batch, and a pageable copy of a 271 M-entry
    // pattern took 0.37 s
    halloc(stage, sizeof(int64_t) * ((size_t)n + 1) + sizeof(int32_t) * (size_t)nnz, "sell pattern staging");
    int64_t* const rp = static_cast<int64_t*>(stage.p);
    int32_t* const col = reinterpret_cast<int32_t*>(rp + n + 1);
    check(hipMemcpyAsync(rp, drp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s), "sell");
    if (nnz) check(hipMemcpyAsync(col, dcol, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, s), "sell");
    check(hipStreamSynchronize(s), "sell");
    std::vector<uint8_t> brk(n + 1, 0);
    if (breaks)
        for (int64_t b : *breaks)
            if (b >= 0 && b <= n) brk[b] = 1;
    S.ucap = ucap;
    S.tmax = tmax;
    S.spatial = spatial;
    auto fail = [&]() {
        S.failed = true;
        S.valid = false;
        return false;
    };
    // Stacked jobs never share a tile, so contiguous runs of whole jobs are built
    // on host threads side by side (each job's columns lie in its own rows) and
    // joined in order: the same tiles as one pass over the rows.  A single matrix
    // of kSellSplitRows rows or more is cut into 16 fixed row ranges (the same on
    // every machine) whose first rows start tiles; a smaller one is one part
    // (LEO2D_SELL_SPLIT_ROWS moves that bound).
    const int64_t split_rows = std::max<int64_t>(16, env_int("LEO2D_SELL_SPLIT_ROWS", 1 << 20));
    std::vector<int> cuts{0};
    bool local = true;
    if (breaks && breaks->size() > 2) {
        const int want = (int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency()));
        const int64_t per = (int64_t)n / want + 1;
        for (int64_t b : *breaks)
            if (b > cuts.back() && b < n && b - cuts.back() >= per) cuts.push_back((int)b);
    } else if (!breaks && n >= split_rows) {
        local = false;
        for (int k = 1; k < 16; ++k) cuts.push_back((int)((int64_t)n * k / 16));
    }
    cuts.push_back(n);
    const int nparts = (int)cuts.size() - 1;
    std::vector<SellPart> parts(nparts);
    auto run = [&](int p) {
        build_sell_part(parts[p], rp, col, brk.data(), cuts[p], cuts[p + 1], ucap, tmax, local, n);
    };
    if (nparts == 1) {
        run(0);
    } else {
        std::vector<std::thread> pool;
        for (int p = 0; p < nparts; ++p) pool.emplace_back(run, p);
        for (auto& t : pool) t.join();
    }
    if (!std::all_of(parts.begin(), parts.end(), [](const SellPart& P) { return P.ok; })) {
        if (nparts == 1) return fail();
        // a part whose rows reach outside it: build the whole matrix as one part
        parts.assign(1, SellPart{});
        build_sell_part(parts[0], rp, col, brk.data(), 0, n, ucap, tmax, true, n);
        if (!parts[0].ok) return fail();
    }
    // Device arrays, filled part by part at each part's offsets; only the tile and
    // group offsets (u0, g0, goff) are shifted on the host.  row0 is also kept on
    // the host for the job starts.
    const int np = (int)parts.size();
    std::vector<size_t> ot(np + 1, 0), ou(np + 1, 0), og(np + 1, 0), oe(np + 1, 0);
    int tile_rows = 0, max_len = 0;
    for (int p = 0; p < np; ++p) {
        const SellPart& P = parts[p];
        ot[p + 1] = ot[p] + P.row0.size();
        ou[p + 1] = ou[p] + P.ucols.size();
        og[p + 1] = og[p] + P.glen.size();
        oe[p + 1] = oe[p] + P.eu.size();
        tile_rows = std::max(tile_rows, P.tile_rows);
        max_len = std::max(max_len, P.max_len);
    }
    const size_t T = ot[np], U = ou[np], G = og[np], E = oe[np];
    dalloc(S.row0, sizeof(int32_t) * (T + 1), "sell tiles");
    dalloc(S.u0, sizeof(int64_t) * (T + 1), "sell tiles");
    dalloc(S.g0, sizeof(int32_t) * (T + 1), "sell tiles");
    dalloc(S.ucols, sizeof(int32_t) * U, "sell tiles");
    dalloc(S.grow, sizeof(int32_t) * 4 * G, "sell groups");
    dalloc(S.glen, sizeof(int32_t) * G, "sell groups");
    dalloc(S.goff, sizeof(int64_t) * G, "sell groups");
    dalloc(S.eu, sizeof(uint16_t) * E, "sell entries");
    dalloc(S.map, sizeof(int64_t) * E, "sell entries");
    auto put = [&](DevBuf& b, size_t off, const auto& v) {
        using T_ = typename std::decay_t<decltype(v)>::value_type;
        if (!v.empty())
            check(hipMemcpyAsync(static_cast<char*>(b.p) + off * sizeof(T_), v.data(), sizeof(T_) * v.size(),
                                 hipMemcpyHostToDevice, s),
                  "sell tiles");
    };
    std::vector<int32_t> row0;
    row0.reserve(T + 1);
    for (int p = 0; p < np; ++p) {
        SellPart& P = parts[p];
        for (auto& x : P.u0) x += (int64_t)ou[p];
        for (auto& x : P.g0) x += (int32_t)og[p];
        for (auto& x : P.goff) x += (int64_t)oe[p];
        row0.insert(row0.end(), P.row0.begin(), P.row0.end());
        put(S.u0, ot[p], P.u0);
        put(S.g0, ot[p], P.g0);
        put(S.ucols, ou[p], P.ucols);
        put(S.grow, 4 * og[p], P.grow);
        put(S.glen, og[p], P.glen);
        put(S.goff, og[p], P.goff);
        put(S.eu, oe[p], P.eu);
        put(S.map, oe[p], P.map);
    }
    row0.push_back(n);
    put(S.row0, 0, row0);
    const std::vector<int64_t> u_end{(int64_t)U};
    const std::vector<int32_t> g_end{(int32_t)G};
    put(S.u0, T, u_end);
    put(S.g0, T, g_end);
    // the parts' host arrays are read by the copies above
    check(hipStreamSynchronize(s), "sell");
    parts.clear();
    const int tile = (int)T;
    S.job_tile0.clear();
    if (breaks) {
        for (int64_t b : *breaks)
            S.job_tile0.push_back((int32_t)(std::lower_bound(row0.begin(), row0.end() - 1, (int32_t)b) - row0.begin()));
        to_device(S.djob_tile0, S.job_tile0.data(), sizeof(int32_t) * S.job_tile0.size(), s, "sell job tiles");
    }
    check(hipStreamSynchronize(s), "sell");
    S.ntiles = tile;
    S.tile_rows = tile_rows;
    S.max_len = max_len;
    S.ngroups = (int)G;
    S.entries = (int64_t)E;
    S.val_valid = S.val32_valid = false;
    S.valid = true;
    S.failed = false;
    S.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

bool build_sell(leo2d_handle* h, const int64_t* drp, const int32_t* dcol, int ucap, int tmax, bool spatial)
{
    return build_sell_tiles(h->sell, h->sell_stage, h->stream, h->n, h->nnz, drp, dcol, ucap, tmax, spatial, nullptr);
}

// the SELL values of a matrix (FP64, or FP32 for the mixed recursion)
const void* sell_values(SellTiles& S, hipStream_t stream, const void* csr_val, int elem_bytes, bool fp32)
{
    DevBuf& buf = fp32 ? S.val32 : S.val;
    bool& ok = fp32 ? S.val32_valid : S.val_valid;
    if (!ok) {
        dalloc(buf, (size_t)elem_bytes * (size_t)S.entries, "sell values");
        check(leo2d::launch_sell_fill((const int64_t*)S.map.p, csr_val, buf.p, elem_bytes, S.entries, stream),
              "sell values");
        ok = true;
    }
    return buf.p;
}

// Morton (Z-order) rank of every orbital by in-plane position on a 1.42 A
// grid, both sheets interleaved: rows that are neighbours in the lattice end
// up close in memory, so a window of rows and the T_j rows it gathers fit in
// one XCD's L2.
std::vector<int32_t> spatial_order(const leo2d::Geom& g)
{
    const int n = g.n;
    double xmin = 1e300, ymin = 1e300;
    for (int i = 0; i < n; ++i) {
        xmin = std::min(xmin, g.i2p[3 * (size_t)i]);
        ymin = std::min(ymin, g.i2p[3 * (size_t)i + 1]);
    }
    const double cell = 1.42;
    auto spread = [](uint64_t v) {
        v &= 0xffffffffULL;
        v = (v | (v << 16)) & 0x0000FFFF0000FFFFULL;
        v = (v | (v << 8)) & 0x00FF00FF00FF00FFULL;
        v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0FULL;
        v = (v | (v << 2)) & 0x3333333333333333ULL;
        v = (v | (v << 1)) & 0x5555555555555555ULL;
        return v;
    };
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; ++i) {
        const uint64_t ix = (uint64_t)std::max(0.0, std::floor((g.i2p[3 * (size_t)i] - xmin) / cell));
        const uint64_t iy = (uint64_t)std::max(0.0, std::floor((g.i2p[3 * (size_t)i + 1] - ymin) / cell));
        key[i] = spread(ix) | (spread(iy) << 1);
    }
    std::vector<int32_t> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    std::vector<int32_t> perm(n);
    for (int q = 0; q < n; ++q) perm[order[q]] = q;
    return perm;
}

// H' = P H P^T on the device (cached until the Hamiltonian changes)
void build_permuted(leo2d_handle* h)
{
    if (h->pcsr_valid) return;
    const int n = h->n;
    hipStream_t s = h->stream;
    DevBuf len;
    dalloc(len, sizeof(int64_t) * n, "permuted lengths");
    dalloc(h->prp, sizeof(int64_t) * (n + 1), "permuted rows");
    check(leo2d::launch_perm_len((const int64_t*)h->rp.p, (const int32_t*)h->iperm.p, n, (int64_t*)len.p, s), "perm");
    check(leo2d::scan_exclusive_i64((const int64_t*)len.p, (int64_t*)h->prp.p, n, s), "perm scan");
    // kCsrPad trailing entries (column 0, value 0) let the scalar-path step
    // kernel load whole chunks of a row without a bounds clamp
    const size_t padded = (size_t)h->nnz + leo2d::kCsrPad;
    dalloc(h->pcol_k, sizeof(int32_t) * padded, "permuted columns");
    const int vb = h->is_complex ? 16 : 8;
    dalloc(h->pval, (size_t)vb * padded, "permuted values");
    check(hipMemsetAsync((int32_t*)h->pcol_k.p + h->nnz, 0, sizeof(int32_t) * leo2d::kCsrPad, s), "perm pad");
    check(hipMemsetAsync((char*)h->pval.p + (size_t)vb * h->nnz, 0, (size_t)vb * leo2d::kCsrPad, s), "perm pad");
    check(leo2d::launch_perm_fill((const int64_t*)h->rp.p, (const int32_t*)h->col.p, h->val.p, vb,
                                  (const int32_t*)h->iperm.p, (const int32_t*)h->perm.p, n, (const int64_t*)h->prp.p,
                                  (int32_t*)h->pcol_k.p, h->pval.p, s),
          "perm fill");
    check(hipStreamSynchronize(s), "perm");
    h->pcsr_valid = true;
    h->pval32_valid = false;
}


// ---- batched jobs --------------------------------------------------------------
// Move every pending job into the stacked arrays (columns and row pointers
// offset by the rows / entries already stacked).
void batch_stack(leo2d_handle* h)
{
    auto& B = h->batch;
    if (B.pending.empty()) return;
    hipStream_t s = h->stream;
    int64_t rows = B.rows, nnz = B.nnz;
    for (auto& q : B.pending) {
        rows += q.n;
        nnz += q.nnz;
    }
    if (rows > 0x7fffffffLL) throw ArgError("batch exceeds 2^31 rows");
    const size_t vb = B.is_complex ? 16 : 8;
    DevBuf rp, col, val;
    dalloc(rp, sizeof(int64_t) * (size_t)(rows + 1), "batch row pointers");
    dalloc(col, sizeof(int32_t) * (size_t)nnz, "batch columns");
    dalloc(val, vb * (size_t)nnz, "batch values");
    if (B.rows > 0) {
        check(hipMemcpyAsync(rp.p, B.rp.p, sizeof(int64_t) * (size_t)(B.rows + 1), hipMemcpyDeviceToDevice, s), "batch");
        check(hipMemcpyAsync(col.p, B.col.p, sizeof(int32_t) * (size_t)B.nnz, hipMemcpyDeviceToDevice, s), "batch");
        check(hipMemcpyAsync(val.p, B.val.p, vb * (size_t)B.nnz, hipMemcpyDeviceToDevice, s), "batch");
    }
    for (auto& q : B.pending) {
        check(leo2d::launch_batch_offset((const int32_t*)q.col.p, q.nnz, (int32_t)B.rows, (int32_t*)col.p + B.nnz,
                                         (const int64_t*)q.rp.p, q.n, B.nnz, (int64_t*)rp.p + B.rows, s),
              "batch offsets");
        if (q.nnz)
            check(hipMemcpyAsync((char*)val.p + vb * (size_t)B.nnz, q.val.p, vb * (size_t)q.nnz, hipMemcpyDeviceToDevice, s),
                  "batch values");
        B.rows += q.n;
        B.nnz += q.nnz;
        B.job_row0.push_back(B.rows);
    }
    check(hipStreamSynchronize(s), "batch stack");
    B.pending.clear();
    B.rp = std::move(rp);
    B.col = std::move(col);
    B.val = std::move(val);
    B.val32_valid = false;
    to_device(B.drow0, B.job_row0.data(), sizeof(int64_t) * B.job_row0.size(), s, "batch job rows");
    check(hipStreamSynchronize(s), "batch job rows");
    B.tile_rows = 0;
    B.sell.valid = B.sell.failed = false;
    B.sell.val_valid = B.sell.val32_valid = false;
    B.tuned.clear();   // shapes were tuned on the previous stack
    B.drop_graph();
}

// Tiles of at most T rows that never cross a job boundary.
void batch_tiles(leo2d_handle* h, int T)
{
    auto& B = h->batch;
    if (B.tile_rows == T) return;
    // the old tile arrays may still be read by queued steps or a captured graph
    check(hipStreamSynchronize(h->stream), "batch tiles");
    B.drop_graph();
    const int P = (int)B.job_row0.size() - 1;
    B.tile_row0.clear();
    B.job_tile0.assign(1, 0);
    for (int p = 0; p < P; ++p) {
        for (int64_t r = B.job_row0[p]; r < B.job_row0[p + 1]; r += T) B.tile_row0.push_back((int32_t)r);
        B.job_tile0.push_back((int32_t)B.tile_row0.size());
    }
    B.tile_row0.push_back((int32_t)B.rows);
    B.ntiles = (int)B.tile_row0.size() - 1;
    to_device(B.dtile_row0, B.tile_row0.data(), sizeof(int32_t) * B.tile_row0.size(), h->stream, "batch tiles");
    to_device(B.djob_tile0, B.job_tile0.data(), sizeof(int32_t) * B.job_tile0.size(), h->stream, "batch tiles");
    check(hipStreamSynchronize(h->stream), "batch tiles");
    B.tile_rows = T;
}

int batch_default_grid(int ntiles)
{
    return std::min(2048, std::max(8, (ntiles + 7) / 8 * 8));
}

}  // namespace

extern "C" {

// ---- deck -------------------------------------------------------------------
int leo2d_deck_parse(const char* text, leo2d_deck** out, char* err, int errlen)
{
    if (!text || !out) return LEO2D_ERR_ARG;
    try {
        auto* d = new leo2d_deck;
        d->d = leo2d::parse_deck(text);
        *out = d;
        return LEO2D_OK;
    } catch (const std::exception& e) {
        if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", e.what());
        return LEO2D_ERR_DECK;
    }
}

void leo2d_deck_free(leo2d_deck* d) { delete d; }

int leo2d_deck_get_int(const leo2d_deck* d, const char* key, int64_t* value)
{
    if (!d || !key || !value) return LEO2D_ERR_ARG;
    auto it = d->d.opts.i.find(key);
    if (it == d->d.opts.i.end()) return LEO2D_ERR_ARG;
    *value = it->second;
    return LEO2D_OK;
}

int leo2d_deck_get_double(const leo2d_deck* d, const char* key, double* value)
{
    if (!d || !key || !value) return LEO2D_ERR_ARG;
    auto it = d->d.opts.d.find(key);
    if (it == d->d.opts.d.end()) return LEO2D_ERR_ARG;
    *value = it->second;
    return LEO2D_OK;
}

int leo2d_deck_get_string(const leo2d_deck* d, const char* key, char* buf, int buflen)
{
    if (!d || !key || !buf || buflen <= 0) return LEO2D_ERR_ARG;
    auto it = d->d.opts.s.find(key);
    if (it == d->d.opts.s.end()) return LEO2D_ERR_ARG;
    std::snprintf(buf, (size_t)buflen, "%s", it->second.c_str());
    return LEO2D_OK;
}

// ---- geometry and jobs --------------------------------------------------------
int leo2d_geom_build(const leo2d_deck* d, leo2d_geom** out)
{
    if (!d || !out) return LEO2D_ERR_ARG;
    try {
        auto* g = new leo2d_geom;
        g->deck = d->d;
        leo2d::setup_supercell(g->deck);
        g->g = leo2d::build_geometry(g->deck);
        *out = g;
        return LEO2D_OK;
    } catch (const leo2d::DeckError&) {
        return LEO2D_ERR_DECK;
    } catch (const std::bad_alloc&) {
        return LEO2D_ERR_ALLOC;
    } catch (const std::exception&) {
        return LEO2D_ERR_DECK;
    }
}

void leo2d_geom_free(leo2d_geom* g) { delete g; }

int leo2d_geom_get_info(const leo2d_geom* g, leo2d_geom_info* info)
{
    if (!g || !info) return LEO2D_ERR_ARG;
    std::memset(info, 0, sizeof(*info));
    info->n = g->g.n;
    info->num_sheets = g->g.num_sheets;
    for (int s = 0; s < g->g.num_sheets; ++s) {
        info->sheet_count[s] = g->g.count[s];
        info->angle[s] = g->g.angle[s];
    }
    info->boundary_condition = g->g.bc;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) info->supercell[a][b] = g->g.gsc[a][b];
    return LEO2D_OK;
}

int leo2d_geom_arrays(const leo2d_geom* g, int32_t* index_to_grid, double* index_to_pos)
{
    if (!g) return LEO2D_ERR_ARG;
    if (index_to_grid) std::memcpy(index_to_grid, g->g.i2g.data(), sizeof(int32_t) * g->g.i2g.size());
    if (index_to_pos) std::memcpy(index_to_pos, g->g.i2p.data(), sizeof(double) * g->g.i2p.size());
    return LEO2D_OK;
}

int leo2d_jobs_count(const leo2d_deck* d, const leo2d_geom* g, int32_t* count)
{
    if (!g || !count) return LEO2D_ERR_ARG;
    (void)d;
    try {
        *count = (int32_t)leo2d::build_jobs(g->deck, g->g).size();
        return LEO2D_OK;
    } catch (const std::exception&) {
        return LEO2D_ERR_DECK;
    }
}

int leo2d_job_get(const leo2d_deck* d, const leo2d_geom* g, int32_t index, leo2d_job* job)
{
    if (!g || !job) return LEO2D_ERR_ARG;
    (void)d;
    try {
        auto jobs = leo2d::build_jobs(g->deck, g->g);
        if (index < 0 || index >= (int32_t)jobs.size()) return LEO2D_ERR_ARG;
        *job = jobs[(size_t)index];
        return LEO2D_OK;
    } catch (const std::exception&) {
        return LEO2D_ERR_DECK;
    }
}

// ---- handle -----------------------------------------------------------------
int leo2d_create(int32_t device, leo2d_handle** out)
{
    if (!out) return LEO2D_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return LEO2D_ERR_HIP;
    if (device < 0 || device >= count) return LEO2D_ERR_ARG;
    auto* h = new leo2d_handle;
    h->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
        delete h;
        return LEO2D_ERR_HIP;
    }
    // this device's address of the static scratch (leo2d_kpm_moments_noalloc)
    if (leo2d::scratch_init(device) != hipSuccess) {
        (void)hipStreamDestroy(h->stream);
        delete h;
        return LEO2D_ERR_HIP;
    }
    *out = h;
    return LEO2D_OK;
}

void leo2d_destroy(leo2d_handle* h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
}

const char* leo2d_last_error(const leo2d_handle* h) { return h ? h->err.c_str() : "null handle"; }

int leo2d_set_geometry(leo2d_handle* h, const leo2d_geom* gin)
{
    if (!h || !gin) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        const leo2d::Geom& g = gin->g;
        if (g.n <= 0) throw ArgError("geometry has no orbitals");
        int mat0 = g.sheets[0].mat;
        for (const auto& sd : g.sheets)
            if (sd.mat != mat0) throw ArgError("sheets with different coupling models are not supported");
        h->geom = g;
        h->has_pairs = false;
        h->has_csr = false;
        hipStream_t s = h->stream;
        for (auto& b : h->grids) b.release();
        h->grids.clear();
        h->grids.resize(g.num_sheets);
        leo2d::GeomDev d{};
        d.n = g.n;
        d.num_sheets = g.num_sheets;
        d.norb = g.norb;
        d.bc = g.bc;
        d.sc_search = g.sc_search;
        d.inter_win = leo2d::inter_window(g.sheets[0].mat);   // hstruct.cpp:305, sheet 1's material
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                d.gsc[a][b] = g.gsc[a][b];
                d.ainv[a][b] = g.ainv[a][b];
                d.lat[a][b] = g.lat[a][b];
            }
        for (int sidx = 0; sidx < g.num_sheets; ++sidx) {
            to_device(h->grids[sidx], g.grid[sidx].data(), sizeof(int32_t) * g.grid[sidx].size(), s, "grid upload");
            leo2d::SheetDev& sd = d.s[sidx];
            const leo2d::SheetData& src = g.sheets[sidx];
            sd.grid = static_cast<const int32_t*>(h->grids[sidx].p);
            sd.H = g.H[sidx];
            sd.W = g.W[sidx];
            sd.base = g.base[sidx];
            sd.count = g.count[sidx];
            sd.min0 = src.min_shape[0];
            sd.min1 = src.min_shape[1];
            sd.max0 = src.max_shape[0];
            sd.max1 = src.max_shape[1];
            sd.mat = src.mat;
            sd.win = leo2d::intra_window(src.mat);   // sheet.cpp:39, 626
            sd.bc = src.boundary_condition;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    sd.sc[a][b] = src.supercell[a][b];
                    sd.st[a][b] = src.stride[a][b];
                }
            sd.angle = g.angle[sidx];
            const double th = -g.angle[sidx];   // Hstruct::findNearest rotates by -angle
            sd.cneg = std::cos(th);
            sd.sneg = std::sin(th);
        }
        to_device(h->i2g, g.i2g.data(), sizeof(int32_t) * g.i2g.size(), s, "index upload");
        to_device(h->i2p, g.i2p.data(), sizeof(double) * g.i2p.size(), s, "position upload");
        to_device(h->lpos, g.lpos.data(), sizeof(double) * g.lpos.size(), s, "position upload");
        d.i2g = static_cast<const int32_t*>(h->i2g.p);
        d.i2p = static_cast<const double*>(h->i2p.p);
        d.lpos = static_cast<const double*>(h->lpos.p);
        h->gdev = d;
        h->perm_host = spatial_order(g);
        std::vector<int32_t> ip(g.n);
        for (int i = 0; i < g.n; ++i) ip[h->perm_host[i]] = i;
        to_device(h->perm, h->perm_host.data(), sizeof(int32_t) * g.n, s, "order upload");
        to_device(h->iperm, ip.data(), sizeof(int32_t) * g.n, s, "order upload");
        check(hipStreamSynchronize(s), "geometry upload");
        h->has_geom = true;
        return LEO2D_OK;
    });
}

int leo2d_pairs_method(leo2d_handle* h, int32_t* staged)
{
    if (!h || !staged) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_pairs) throw StateError("no pair lists: call leo2d_pairs_build");
        *staged = h->pairs_staged ? 1 : 0;
        return LEO2D_OK;
    });
}

int leo2d_pairs_build(leo2d_handle* h, int64_t* n_intra, int64_t* n_inter, int64_t* n_slots)
{
    if (!h) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_geom) throw StateError("leo2d_pairs_build before leo2d_set_geometry");
        const int n = h->gdev.n;
        hipStream_t s = h->stream;
        dalloc(h->cnt_intra, sizeof(int32_t) * n, "pair counts");
        dalloc(h->cnt_tmp, sizeof(int64_t) * n, "pair counts");
        dalloc(h->pair_rp, sizeof(int64_t) * (n + 1), "pair offsets");
        check(leo2d::pairs_count(h->gdev, (int32_t*)h->cnt_intra.p, (int64_t*)h->cnt_tmp.p, (int64_t*)h->pair_rp.p, s),
              "pair count");
        int64_t total = 0;
        check(hipMemcpyAsync(&total, (int64_t*)h->pair_rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s), "pair total");
        std::vector<int32_t> ci(n);
        std::vector<int64_t> ct(n);
        check(hipMemcpyAsync(ci.data(), h->cnt_intra.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s), "pair counts");
        check(hipMemcpyAsync(ct.data(), h->cnt_tmp.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s), "pair counts");
        check(hipStreamSynchronize(s), "pair count");
        int64_t ni = 0, max_intra = 0, max_inter = 0;
        for (int i = 0; i < n; ++i) {
            ni += ci[i];
            max_intra = std::max<int64_t>(max_intra, ci[i]);
            max_inter = std::max<int64_t>(max_inter, ct[i] - ci[i]);
        }
        dalloc(h->pcol, sizeof(int32_t) * (size_t)total, "pair columns");
        dalloc(h->psc, sizeof(int32_t) * (size_t)total, "pair images");
        dalloc(h->nslots, sizeof(int64_t) * n, "slot counts");
        dalloc(h->slot_rp, sizeof(int64_t) * (n + 1), "slot offsets");
        // lists staged in LDS, one wavefront per row, where every row's lists fit; else
        // one thread per row through a scratch copy of the pairs
        h->pairs_staged = env_int("LEO2D_PAIRS_STAGED", 1) != 0 &&
                          leo2d::pairs_staged_fit(h->gdev, (int)std::min<int64_t>(max_intra, INT32_MAX),
                                                  (int)std::min<int64_t>(max_inter, INT32_MAX));
        if (h->pairs_staged) {
            check(leo2d::pairs_write_staged(h->gdev, (const int64_t*)h->pair_rp.p, (int32_t*)h->pcol.p,
                                            (int32_t*)h->psc.p, (int64_t*)h->nslots.p, (int64_t*)h->slot_rp.p, s),
                  "pair write");
        } else {
            DevBuf tcol, tsc;   // the pairs as found, before the per-image runs are merged
            dalloc(tcol, sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1), "pair scratch");
            dalloc(tsc, sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1), "pair scratch");
            check(leo2d::pairs_write(h->gdev, (const int64_t*)h->pair_rp.p, (const int32_t*)h->cnt_intra.p,
                                     (int32_t*)tcol.p, (int32_t*)tsc.p, (int32_t*)h->pcol.p, (int32_t*)h->psc.p,
                                     (int64_t*)h->nslots.p, (int64_t*)h->slot_rp.p, s),
                  "pair write");
            check(hipStreamSynchronize(s), "pair write");
        }
        int64_t slots = 0;
        check(hipMemcpyAsync(&slots, (int64_t*)h->slot_rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s), "slot total");
        check(hipStreamSynchronize(s), "pair write");
        h->cnt_tmp.release();
        h->nslots.release();
        h->n_pairs = total;
        h->n_intra = ni;
        h->n_slots = slots;
        h->has_pairs = true;
        if (n_intra) *n_intra = ni;
        if (n_inter) *n_inter = total - ni;
        if (n_slots) *n_slots = slots;
        return LEO2D_OK;
    });
}

int leo2d_pairs_copy(leo2d_handle* h, int32_t which, int32_t* ij, int32_t* sc, double* t)
{
    if (!h || (which != 0 && which != 1)) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_pairs) throw StateError("no pair lists: call leo2d_pairs_build");
        const int n = h->gdev.n;
        std::vector<int64_t> rp(n + 1);
        std::vector<int32_t> ci(n), pc(h->n_pairs), ps(h->n_pairs);
        hipStream_t s = h->stream;
        check(hipMemcpyAsync(rp.data(), h->pair_rp.p, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s), "copy");
        check(hipMemcpyAsync(ci.data(), h->cnt_intra.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s), "copy");
        if (h->n_pairs) {
            check(hipMemcpyAsync(pc.data(), h->pcol.p, sizeof(int32_t) * h->n_pairs, hipMemcpyDeviceToHost, s), "copy");
            check(hipMemcpyAsync(ps.data(), h->psc.p, sizeof(int32_t) * h->n_pairs, hipMemcpyDeviceToHost, s), "copy");
        }
        check(hipStreamSynchronize(s), "copy");
        const leo2d::Geom& g = h->geom;
        int64_t w = 0;
        for (int r = 0; r < n; ++r) {
            const int64_t b = which == 0 ? rp[r] : rp[r] + ci[r];
            const int64_t e = which == 0 ? rp[r] + ci[r] : rp[r + 1];
            for (int64_t q = b; q < e; ++q, ++w) {
                const int dx = (int)(int16_t)(ps[q] & 0xffff), dy = (int)(int16_t)((uint32_t)ps[q] >> 16);
                if (ij) { ij[2 * w] = r; ij[2 * w + 1] = pc[q]; }
                if (sc) { sc[2 * w] = dx; sc[2 * w + 1] = dy; }
                if (t && which == 0) {
                    // tabulated intralayer hopping of the pair (graphene.cpp:13-89)
                    const int s0 = g.i2g[4 * (size_t)r + 3];
                    const leo2d::SheetData& sd = g.sheets[s0];
                    int gd0 = g.i2g[4 * (size_t)pc[q]] - g.i2g[4 * (size_t)r];
                    int gd1 = g.i2g[4 * (size_t)pc[q] + 1] - g.i2g[4 * (size_t)r + 1];
                    if (sd.boundary_condition == 1) {
                        gd0 = gd0 - dx * sd.stride[0][0] - dy * sd.stride[1][0];
                        gd1 = gd1 - dx * sd.stride[0][1] - dy * sd.stride[1][1];
                    }
                    const int o1 = g.i2g[4 * (size_t)r + 2], o2 = g.i2g[4 * (size_t)pc[q] + 2];
                    int h0 = 2 * gd0 + gd1;
                    const int h1 = -1 * gd0 + gd1;
                    if (o2 == 0 && o1 == 1) h0 -= 1;
                    else if (o2 == 1 && o1 == 0) h0 += 1;
                    const int rr = h0 * (h0 + h1) + h1 * h1;
                    static const double tab[17] = {0.3504, -2.8922, 0, 0.2425, -0.2656, 0, 0, 0.0235, 0, 0.0524,
                                                   0, 0, -0.0209, -0.0148, 0, 0, -0.0211};
                    double v = (rr >= 0 && rr <= 16) ? tab[rr] : 0.0;
                    if (sd.mat == leo2d::MAT_STRAINED_GRAPHENE) {
                        // strained_graphene.cpp:16-91
                        v = rr == 0 ? 0.7833 : rr == 1 ? -2.8219 : rr == 3 ? 0.2543 : rr == 4 ? -0.1803 : 0.0;
                    } else if (sd.mat != leo2d::MAT_GRAPHENE) {
                        v = rr == 1 ? -2.7 : 0.0;
                    }
                    t[w] = v;
                }
            }
        }
        return LEO2D_OK;
    });
}

namespace {
int assemble_impl(leo2d_handle* h, const leo2d_job* job, int32_t pattern, int64_t* nnz_out, bool velocity)
{
    if (!h || !job || (pattern != LEO2D_PATTERN_REFERENCE && pattern != LEO2D_PATTERN_PRUNED)) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_geom) throw StateError("leo2d_assemble before leo2d_set_geometry");
        if (!(job->energy_rescale != 0.0) || !std::isfinite(job->energy_rescale))
            throw ArgError("energy_rescale must be finite and non-zero");
        if (job->sheet_strain_on && h->geom.bc == 1) throw ArgError("per-sheet strain needs open boundaries");
        if (job->strain_type != 0 && job->strain_type != 1 && job->strain_type != 6)
            throw ArgError("strain_type must be 0, 1 or 6");
        if (job->strain_type != 0 && h->geom.bc != 1) throw ArgError("the relaxation field needs a periodic supercell");
        if (job->strain_type == 1 && h->relax_max_k < 0) throw StateError("strain_type 1 before leo2d_set_relaxation");
        if (!h->has_pairs) {
            int rc = leo2d_pairs_build(h, nullptr, nullptr, nullptr);
            if (rc != LEO2D_OK) return rc;
        }
        const int n = h->gdev.n;
        const int cpx = job->complex_matrix ? 1 : 0;
        hipStream_t s = h->stream;
        leo2d::JobDev jd = job_dev(h->geom, *job);
        if (!h->i2pos.p || h->i2pos.bytes < sizeof(double) * 3 * (size_t)n)
            dalloc(h->i2pos, sizeof(double) * 3 * (size_t)n, "job positions");
        if (job->strain_type == 1) {
            jd.relax = (const double*)h->relax.p;
            jd.relax_max_k = h->relax_max_k;
        }
        check(leo2d::job_positions(h->gdev, jd, (double*)h->i2pos.p, s), "job positions");
        const double* lt = nullptr;
        if (job->strain_type != 0) {
            if (!h->ltheta.p || h->ltheta.bytes < sizeof(double) * (size_t)std::max(n, 1))
                dalloc(h->ltheta, sizeof(double) * (size_t)std::max(n, 1), "local twist");
            check(leo2d::local_theta(h->gdev, (const double*)h->i2pos.p, (double*)h->ltheta.p, s), "local twist");
            lt = (const double*)h->ltheta.p;
        }
        DevBuf scol, sval, skeep, sdx, sdy;
        dalloc(scol, sizeof(int32_t) * (size_t)h->n_slots, "slot columns");
        dalloc(sval, sizeof(double) * (cpx ? 2 : 1) * (size_t)h->n_slots, "slot values");
        dalloc(skeep, (size_t)h->n_slots + 1, "slot flags");
        if (velocity) {
            dalloc(sdx, sizeof(double) * (cpx ? 2 : 1) * (size_t)h->n_slots, "x-velocity slots");
            dalloc(sdy, sizeof(double) * (cpx ? 2 : 1) * (size_t)h->n_slots, "y-velocity slots");
        }
        check(leo2d::fill_slots(h->gdev, jd, (const double*)h->i2pos.p, (const int64_t*)h->pair_rp.p,
                                (const int32_t*)h->cnt_intra.p, (const int32_t*)h->pcol.p, (const int32_t*)h->psc.p,
                                (const int64_t*)h->slot_rp.p, (int32_t*)scol.p, (double*)sval.p, (uint8_t*)skeep.p,
                                velocity ? (double*)sdx.p : nullptr, velocity ? (double*)sdy.p : nullptr, lt, s),
              "hamiltonian fill");
        h->rp.release();
        h->col.release();
        h->val.release();
        h->val_dx.release();
        h->val_dy.release();
        if (pattern == LEO2D_PATTERN_REFERENCE) {
            dalloc(h->rp, sizeof(int64_t) * (n + 1), "row pointers");
            check(hipMemcpyAsync(h->rp.p, h->slot_rp.p, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice, s), "rp");
            h->col = std::move(scol);
            h->val = std::move(sval);
            skeep.release();
            set_csr_owned(h, n, h->n_slots, cpx);
            if (velocity) {
                h->val_dx = std::move(sdx);
                h->val_dy = std::move(sdy);
            }
        } else {
            DevBuf cnt;
            dalloc(cnt, sizeof(int64_t) * n, "prune counts");
            dalloc(h->rp, sizeof(int64_t) * (n + 1), "row pointers");
            check(leo2d::prune_count(n, (const int64_t*)h->slot_rp.p, (const uint8_t*)skeep.p, (int64_t*)cnt.p,
                                     (int64_t*)h->rp.p, s),
                  "prune count");
            int64_t kept = 0;
            check(hipMemcpyAsync(&kept, (int64_t*)h->rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s), "prune");
            check(hipStreamSynchronize(s), "prune");
            dalloc(h->col, sizeof(int32_t) * (size_t)kept, "columns");
            dalloc(h->val, sizeof(double) * (cpx ? 2 : 1) * (size_t)kept, "values");
            check(leo2d::prune_compact(n, cpx, (const int64_t*)h->slot_rp.p, (const uint8_t*)skeep.p,
                                       (const int32_t*)scol.p, (const double*)sval.p, (const int64_t*)h->rp.p,
                                       (int32_t*)h->col.p, (double*)h->val.p, s),
                  "prune compact");
            if (velocity) {
                // the same compaction for the velocity slots (columns rewritten identically)
                DevBuf* dst[2] = {&h->val_dx, &h->val_dy};
                const DevBuf* src[2] = {&sdx, &sdy};
                for (int c = 0; c < 2; ++c) {
                    dalloc(*dst[c], sizeof(double) * (cpx ? 2 : 1) * (size_t)std::max<int64_t>(kept, 1),
                           "velocity values");
                    check(leo2d::prune_compact(n, cpx, (const int64_t*)h->slot_rp.p, (const uint8_t*)skeep.p,
                                               (const int32_t*)scol.p, (const double*)src[c]->p,
                                               (const int64_t*)h->rp.p, (int32_t*)h->col.p, (double*)dst[c]->p, s),
                          "prune compact velocity");
                }
            }
            check(hipStreamSynchronize(s), "prune compact");
            cnt.release();
            scol.release();
            sval.release();
            skeep.release();
            sdx.release();
            sdy.release();
            set_csr_owned(h, n, kept, cpx);
        }
        check(hipStreamSynchronize(s), "assemble");
        h->csr_geom_order = true;
        h->has_dx = velocity;
        if (nnz_out) *nnz_out = h->nnz;
        return LEO2D_OK;
    });
}
}  // namespace

int leo2d_assemble(leo2d_handle* h, const leo2d_job* job, int32_t pattern, int64_t* nnz_out)
{
    return assemble_impl(h, job, pattern, nnz_out, false);
}

int leo2d_assemble_velocity(leo2d_handle* h, const leo2d_job* job, int32_t pattern, int64_t* nnz_out)
{
    return assemble_impl(h, job, pattern, nnz_out, true);
}

int leo2d_set_relaxation(leo2d_handle* h, int32_t max_k, const double* coeffs)
{
    if (!h) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (max_k < 0) {
            h->relax.release();
            h->relax_max_k = -1;
            return LEO2D_OK;
        }
        if (!coeffs) throw ArgError("relaxation coefficients missing");
        if (max_k > 64) throw ArgError("relaxation max_k above 64");
        const size_t count = 3 * ((size_t)max_k * (max_k + 1) / 2 + 1);
        for (size_t q = 0; q < count; ++q)
            if (!std::isfinite(coeffs[q])) throw NanError("relaxation coefficient is not finite");
        check(hipStreamSynchronize(h->stream), "relaxation");
        dalloc(h->relax, sizeof(double) * count, "relaxation coefficients");
        check(hipMemcpy(h->relax.p, coeffs, sizeof(double) * count, hipMemcpyHostToDevice), "relaxation upload");
        h->relax_max_k = max_k;
        return LEO2D_OK;
    });
}

int leo2d_read_relaxation(const char* path, int32_t* max_k, double* coeffs, int64_t* count)
{
    // StrainCalc::loadFourierConfigFile (strain.cpp:258-284): max_k, then the u_x, u_y and
    // u_z rows of max_k (max_k + 1) / 2 + 1 values each, read as whitespace-separated numbers
    if (!path || !max_k || !count) return LEO2D_ERR_ARG;
    std::ifstream f(path);
    if (!f) return LEO2D_ERR_DECK;
    long long mk = -1;
    if (!(f >> mk) || mk < 0 || mk > 64) return LEO2D_ERR_DECK;
    const int64_t n = 3 * (mk * (mk + 1) / 2 + 1);
    *max_k = (int32_t)mk;
    *count = n;
    if (!coeffs) return LEO2D_OK;
    for (int64_t q = 0; q < n; ++q)
        if (!(f >> coeffs[q])) return LEO2D_ERR_DECK;
    return LEO2D_OK;
}

int leo2d_velocity_copy(leo2d_handle* h, double* val_dx, double* val_dy)
{
    if (!h || (!val_dx && !val_dy)) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_csr || !h->has_dx) throw StateError("no velocity matrix: use leo2d_assemble_velocity");
        const size_t bytes = sizeof(double) * (h->is_complex ? 2 : 1) * (size_t)h->nnz;
        if (h->nnz && val_dx) check(hipMemcpy(val_dx, h->val_dx.p, bytes, hipMemcpyDeviceToHost), "velocity copy");
        if (h->nnz && val_dy) check(hipMemcpy(val_dy, h->val_dy.p, bytes, hipMemcpyDeviceToHost), "velocity copy");
        return LEO2D_OK;
    });
}

int leo2d_csr_upload(leo2d_handle* h, int32_t n, int64_t nnz, const int64_t* row_ptr, const int32_t* col,
                     const double* val, int32_t is_complex)
{
    if (!h || n <= 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col || !val))) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (row_ptr[0] != 0 || row_ptr[n] != nnz) throw ArgError("row_ptr must start at 0 and end at nnz");
        for (int i = 0; i < n; ++i)
            if (row_ptr[i + 1] < row_ptr[i]) throw ArgError("row_ptr must be non-decreasing");
        for (int64_t q = 0; q < nnz; ++q)
            if (col[q] < 0 || col[q] >= n) throw ArgError("column index out of range");
        const size_t w = is_complex ? 2 : 1;
        for (size_t q = 0; q < w * (size_t)nnz; ++q)
            if (!std::isfinite(val[q])) throw NanError("non-finite matrix value at entry " + std::to_string(q / w));
        hipStream_t s = h->stream;
        to_device(h->rp, row_ptr, sizeof(int64_t) * (n + 1), s, "row pointers");
        to_device(h->col, col, sizeof(int32_t) * (size_t)nnz, s, "columns");
        to_device(h->val, val, sizeof(double) * w * (size_t)nnz, s, "values");
        check(hipStreamSynchronize(s), "upload");
        set_csr_owned(h, n, nnz, is_complex ? 1 : 0);
        h->csr_geom_order = false;
        return LEO2D_OK;
    });
}

int leo2d_csr_info(leo2d_handle* h, int32_t* n, int64_t* nnz, int32_t* is_complex)
{
    if (!h) return LEO2D_ERR_ARG;
    if (!h->has_csr) {
        h->err = "no Hamiltonian on the device";
        return LEO2D_ERR_STATE;
    }
    if (n) *n = h->n;
    if (nnz) *nnz = h->nnz;
    if (is_complex) *is_complex = h->is_complex;
    return LEO2D_OK;
}

int leo2d_csr_copy(leo2d_handle* h, int64_t* row_ptr, int32_t* col, double* val)
{
    if (!h) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_csr) throw StateError("no Hamiltonian on the device");
        hipStream_t s = h->stream;
        if (row_ptr) check(hipMemcpyAsync(row_ptr, h->rp.p, sizeof(int64_t) * (h->n + 1), hipMemcpyDeviceToHost, s), "copy");
        if (col && h->nnz) check(hipMemcpyAsync(col, h->col.p, sizeof(int32_t) * h->nnz, hipMemcpyDeviceToHost, s), "copy");
        if (val && h->nnz)
            check(hipMemcpyAsync(val, h->val.p, sizeof(double) * (h->is_complex ? 2 : 1) * h->nnz, hipMemcpyDeviceToHost, s),
                  "copy");
        check(hipStreamSynchronize(s), "copy");
        return LEO2D_OK;
    });
}

int leo2d_spectral_bounds(leo2d_handle* h, double* lo, double* hi)
{
    if (!h || !lo || !hi) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_csr) throw StateError("no Hamiltonian on the device");
        const int blocks = 1024;
        DevBuf buf;
        dalloc(buf, sizeof(double) * 2 * blocks, "bounds");
        check(leo2d::launch_gershgorin((const int64_t*)h->rp.p, (const int32_t*)h->col.p, h->val.p, h->is_complex, h->n,
                                       (double*)buf.p, blocks, h->stream),
              "gershgorin");
        std::vector<double> v(2 * blocks);
        check(hipMemcpyAsync(v.data(), buf.p, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost, h->stream), "bounds");
        check(hipStreamSynchronize(h->stream), "bounds");
        double l = 1e300, u = -1e300;
        for (int b = 0; b < blocks; ++b) {
            l = std::min(l, v[2 * b]);
            u = std::max(u, v[2 * b + 1]);
        }
        *lo = l;
        *hi = u;
        return LEO2D_OK;
    });
}

// ---- KPM --------------------------------------------------------------------
namespace {

// workgroups of a step launch when the caller leaves the grid to the library.
// R >= 64: four workgroups per CU (one wavefront per row); fewer rows in flight
// keep each XCD's gathered rows in its L2 (measured best on the 1.05 deg 8x8
// cell: 1.66 ms per FP64 step at 1024 vs 1.79 ms at 2048)
int auto_grid(int n, int B)
{
    const int rows_per_iter = std::max(1, 256 / B);   // 4 wavefronts x 64/B rows
    const int want = (n + rows_per_iter - 1) / rows_per_iter;
    return std::min(B >= 64 ? 1024 : 2048, std::max(8, (want + 7) / 8 * 8));
}

// moments of a spectrum inside [-1, 1] are finite and satisfy |mu_n| <= mu_0
int check_moments(leo2d_handle* h, const double* mu, int R, int M)
{
    for (int r = 0; r < R; ++r) {
        const double* m = mu + (size_t)r * M;
        for (int j = 0; j < M; ++j)
            if (!std::isfinite(m[j])) {
                h->err = "non-finite moment (vector " + std::to_string(r) + ", n = " + std::to_string(j) + ")";
                return LEO2D_ERR_NAN;
            }
        const double lim = std::fabs(m[0]) * (1.0 + 1e-6) + 1e-12;
        for (int j = 0; j < M; ++j)
            if (std::fabs(m[j]) > lim) {
                h->err = "|mu_" + std::to_string(j) + "| > mu_0: rescaled spectrum leaves [-1, 1]";
                return LEO2D_ERR_SPECTRUM;
            }
    }
    return LEO2D_OK;
}

// the arguments leo2d_kpm_moments and its allocation-free variant check alike
void check_kpm_params(const leo2d_handle* h, const leo2d_kpm_params* p)
{
    if (!h->has_csr) throw StateError("KPM moments before a Hamiltonian is installed");
    const int M = p->M, R = p->R, n = h->n;
    if (M < 4 || M % 4 != 0) throw ArgError("M must be a positive multiple of 4 (POLY_ORDER % 4, main.cpp:914)");
    if (R < 1) throw ArgError("R must be >= 1");
    if (p->start < 0 || p->start > 2) throw ArgError("unknown start mode");
    if (p->precision != LEO2D_PREC_FP64 && p->precision != LEO2D_PREC_MIXED) throw ArgError("unknown precision");
    if (p->scheme != LEO2D_SCHEME_REFERENCE && p->scheme != LEO2D_SCHEME_DOUBLING) throw ArgError("unknown scheme");
    if (p->start == LEO2D_START_TARGETS) {
        if (!p->targets) throw ArgError("targets missing");
        for (int r = 0; r < R; ++r)
            if (p->targets[r] < 0 || p->targets[r] >= n) throw ArgError("target index out of range");
    }
    if (p->start == LEO2D_START_VECTORS && !p->vectors) throw ArgError("vectors missing");
    const int B = p->block ? p->block : std::min(64, next_pow2(R));
    if (B < 1 || B > 128 || (B & (B - 1))) throw ArgError("block must be a power of two <= 128");
}

}  // namespace

int leo2d_kpm_moments(leo2d_handle* h, const leo2d_kpm_params* p, double* mu)
{
    if (!h || !p || !mu) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        check_kpm_params(h, p);
        const int M = p->M, R = p->R, n = h->n;
        const int B = p->block ? p->block : std::min(64, next_pow2(R));
        const bool mixed = p->precision == LEO2D_PREC_MIXED;
        const bool real_vec = !h->is_complex &&
                              (p->start == LEO2D_START_TARGETS ||
                               (p->start == LEO2D_START_VECTORS && !p->vectors_complex));
        int mat_type, vec_type;
        if (!mixed) {
            mat_type = h->is_complex ? leo2d::MAT_CPX_F64 : leo2d::MAT_REAL_F64;
            vec_type = real_vec ? leo2d::VEC_REAL_F64 : leo2d::VEC_CPX_F64;
        } else {
            mat_type = h->is_complex ? leo2d::MAT_CPX_F32 : leo2d::MAT_REAL_F32;
            vec_type = real_vec ? leo2d::VEC_REAL_F32 : leo2d::VEC_CPX_F32;
        }
        hipStream_t s = h->stream;
        const bool spatial = !p->natural_order && h->csr_geom_order && h->has_geom && h->gdev.n == n;
        const int32_t* iperm = nullptr;
        const int64_t* drp = (const int64_t*)h->rp.p;
        const int32_t* dcol = (const int32_t*)h->col.p;
        const void* dval = h->val.p;
        DevBuf* v32 = &h->val32;
        bool* v32_valid = &h->val32_valid;
        if (spatial) {
            build_permuted(h);
            iperm = (const int32_t*)h->iperm.p;
            drp = (const int64_t*)h->prp.p;
            dcol = (const int32_t*)h->pcol_k.p;
            dval = h->pval.p;
            v32 = &h->pval32;
            v32_valid = &h->pval32_valid;
        }
        if (mixed) {
            const size_t count = (size_t)h->nnz * (h->is_complex ? 2 : 1);
            if (!*v32_valid) {
                const size_t pad = (size_t)leo2d::kCsrPad * (h->is_complex ? 2 : 1);
                dalloc(*v32, sizeof(float) * (count + pad), "fp32 matrix");
                check(hipMemsetAsync((float*)v32->p + count, 0, sizeof(float) * pad, s), "fp32 pad");
                check(leo2d::launch_convert_f64_to_f32((const double*)dval, (float*)v32->p, (int64_t)count, s),
                      "fp32 matrix");
                *v32_valid = true;
            }
            dval = v32->p;
        }
        // scheme on the device (see kernels.h)
        int dev_scheme;
        if (p->scheme == LEO2D_SCHEME_DOUBLING) dev_scheme = 1;
        else dev_scheme = p->start == LEO2D_START_TARGETS ? 0 : 2;
        const bool keep_v = dev_scheme == 2;
        const int G = p->grid > 0 ? p->grid : auto_grid(n, B);
        if (G % 8 != 0) throw ArgError("grid must be a multiple of 8");
        const size_t vb = leo2d::vec_bytes(vec_type);
        DevBuf A, Bf, V, part, dmu, dtg, host_v;
        dalloc(A, vb * (size_t)n * B, "vector block");
        dalloc(Bf, vb * (size_t)n * B, "vector block");
        if (keep_v) dalloc(V, vb * (size_t)n * B, "start vectors");
        dalloc(part, sizeof(double) * 3 * (size_t)G * B, "partials");
        dalloc(dmu, sizeof(double) * (size_t)M * B, "moments");
        dalloc(dtg, sizeof(int32_t) * B, "targets");
        std::vector<double> hv;
        double rec_ms = 0.0, step_ms = 0.0;
        int kernel_used = -1;
        int64_t steps = 0;
        hipEvent_t e0 = event_at(h, 0), e1 = event_at(h, 1);
        for (int b0 = 0; b0 < R; b0 += B) {
            const int Rb = std::min(B, R - b0);
            check(hipMemsetAsync(dmu.p, 0, sizeof(double) * (size_t)M * B, s), "moments");
            if (p->start == LEO2D_START_TARGETS) {
                std::vector<int32_t> tg(B);
                for (int r = 0; r < B; ++r) {
                    const int t = p->targets[b0 + std::min(r, Rb - 1)];
                    tg[r] = spatial ? h->perm_host[t] : t;
                }
                check(hipMemcpyAsync(dtg.p, tg.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, s), "targets");
                check(hipStreamSynchronize(s), "targets");
                check(leo2d::launch_init_targets(A.p, vec_type, n, B, (const int32_t*)dtg.p, s), "init");
            } else if (p->start == LEO2D_START_RANDOM_PHASE) {
                if (real_vec) throw ArgError("random phases need complex vectors");
                check(leo2d::launch_init_random_phase_perm(A.p, vec_type, n, B, p->seed, b0, iperm, s), "init");
            } else {
                const int w = p->vectors_complex ? 2 : 1;
                hv.assign((size_t)w * n * B, 0.0);
                std::memcpy(hv.data(), p->vectors + (size_t)w * n * b0, sizeof(double) * (size_t)w * n * Rb);
                dalloc(host_v, sizeof(double) * hv.size(), "vector upload");
                check(hipMemcpyAsync(host_v.p, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice, s), "upload");
                check(leo2d::launch_init_from_host_layout_perm(A.p, vec_type, (const double*)host_v.p,
                                                               p->vectors_complex, n, B, iperm, s),
                      "init");
            }
            if (keep_v) check(hipMemcpyAsync(V.p, A.p, vb * (size_t)n * B, hipMemcpyDeviceToDevice, s), "copy start");
            check(hipStreamSynchronize(s), "init");
            leo2d::StepArgs sa{};
            sa.row_ptr = drp;
            sa.col = dcol;
            sa.val = dval;
            sa.n = n;
            sa.R = B;
            sa.part = (double*)part.p;
            sa.grid = G;
            sa.v = keep_v ? V.p : nullptr;
            sa.variant = kpm_variant(mixed, spatial, B, n);
            sa.xcd_rows = kpm_sched() == 1 ? (n + 7) / 8 : 0;
            leo2d::SellStepArgs ss{};
            const bool sell_alternate = env_int("LEO2D_SELL_ALTERNATE", 1) != 0;
            if (sa.variant == kVariantSell) {
                // one workgroup per CU with two LDS buffers of kSellUcap x 256 B; tiles of at most
                // tmax rows whose union fits ucap staged rows
                const int ucap = std::max(32, std::min(leo2d::kSellUcap, env_int("LEO2D_SELL_UCAP", leo2d::kSellUcap)));
                // a copy wavefront beside seven compute wavefronts (tiles of at most 56 rows), or
                // eight wavefronts that each issue their share of the copies
                const bool copy_wave = env_int("LEO2D_SELL_CW", 1) != 0;
                const int tlim = copy_wave ? leo2d::kSellCwRows : leo2d::kSellMaxRows;
                const int tmax = std::max(4, std::min(tlim, env_int("LEO2D_SELL_TMAX", tlim)));
                if (!spatial || !leo2d::sell_supported(mat_type, vec_type, B, copy_wave) ||
                    !build_sell(h, drp, dcol, ucap, tmax, spatial)) {
                    sa.variant = B == 16 || B == 32 ? 14 : (mixed ? 15 : 1);
                } else {
                    auto& S = h->sell;
                    const int hb = mat_type == leo2d::MAT_CPX_F64 ? 16
                                   : (mat_type == leo2d::MAT_REAL_F32 ? 4 : 8);
                    ss.ev = sell_values(S, h->stream, dval, hb, mixed);
                    int dev_id = 0, cus = 256;
                    if (hipGetDevice(&dev_id) == hipSuccess)
                        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id);
                    ss.grid = std::max(8, std::min(G, cus) / 8 * 8);
                    sa.grid = ss.grid;   // the reduce reads this many partials
                    ss.n = n;
                    ss.R = B;
                    ss.part = (double*)part.p;
                    ss.v = keep_v ? V.p : nullptr;
                    ss.ntiles = S.ntiles;
                    ss.tile_row0 = (const int32_t*)S.row0.p;
                    ss.tile_u0 = (const int64_t*)S.u0.p;
                    ss.ucols = (const int32_t*)S.ucols.p;
                    ss.tile_g0 = (const int32_t*)S.g0.p;
                    ss.grow = (const int32_t*)S.grow.p;
                    ss.glen = (const int32_t*)S.glen.p;
                    ss.goff = (const int64_t*)S.goff.p;
                    ss.eu = (const uint16_t*)S.eu.p;
                    ss.tile_rows = S.tile_rows;
                    ss.max_len = S.max_len;
                    ss.sliced = env_int("LEO2D_SELL_SLICED", 0);
                    ss.ucap = S.ucap;
                    ss.tile_part = 0;
                    ss.copy_wave = copy_wave ? 1 : 0;
                }
            }
            if (sa.variant == 14 && B != 16 && B != 32 && B != 64 && B != 128) sa.variant = 1;
            if (sa.variant == 15 && (!spatial || B != 64 || vec_type != leo2d::VEC_CPX_F32)) sa.variant = 1;
            if (B < 64 && sa.variant != 14 && sa.variant != kVariantSell) sa.variant = 0;
            kernel_used = sa.variant == kVariantSell && ss.copy_wave ? kVariantSell + 1 : sa.variant;
            leo2d::ReduceArgs ra{};
            ra.part = (const double*)part.p;
            ra.grid = sa.grid;
            ra.R = B;
            ra.M = M;
            ra.scheme = dev_scheme;
            ra.targets = (const int32_t*)dtg.p;
            ra.vec_type = vec_type;
            ra.mu = (double*)dmu.p;
            check(hipEventRecord(e0, s), "event");
            for (int j = 0; j < M / 2; ++j) {
                void* cur = (j % 2 == 0) ? A.p : Bf.p;
                void* other = (j % 2 == 0) ? Bf.p : A.p;
                sa.cur = cur;
                sa.prev = other;
                sa.out = other;
                sa.alpha = j == 0 ? 1.0 : 2.0;
                sa.beta = j == 0 ? 0 : 1;
                if (p->time_steps) check(hipEventRecord(event_at(h, 2 + 2 * (size_t)j), s), "event");
                if (sa.variant == kVariantSell) {
                    ss.cur = sa.cur;
                    ss.prev = sa.prev;
                    ss.out = sa.out;
                    ss.alpha = sa.alpha;
                    ss.beta = sa.beta;
                    ss.reverse = sell_alternate && (j & 1);
                    check(leo2d::launch_kpm_step_sell(mat_type, vec_type, ss, s), "kpm step");
                } else {
                    check(leo2d::launch_kpm_step(mat_type, vec_type, sa, s), "kpm step");
                }
                if (p->time_steps) check(hipEventRecord(event_at(h, 3 + 2 * (size_t)j), s), "event");
                ra.j = j;
                ra.cur = cur;
                ra.out = other;
                check(leo2d::launch_kpm_reduce(ra, s), "kpm reduce");
            }
            check(hipEventRecord(e1, s), "event");
            check(hipStreamSynchronize(s), "kpm recursion");
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, e0, e1), "event");
            rec_ms += ms;
            if (p->time_steps) {
                for (int j = 0; j < M / 2; ++j) {
                    float sm = 0.f;
                    check(hipEventElapsedTime(&sm, event_at(h, 2 + 2 * (size_t)j), event_at(h, 3 + 2 * (size_t)j)), "event");
                    step_ms += sm;
                }
            }
            steps += M / 2;
            std::vector<double> hm((size_t)M * B);
            check(hipMemcpy(hm.data(), dmu.p, sizeof(double) * (size_t)M * B, hipMemcpyDeviceToHost), "moments");
            std::memcpy(mu + (size_t)M * b0, hm.data(), sizeof(double) * (size_t)M * Rb);
        }
        h->timing.recursion_ms = rec_ms;
        h->timing.step_ms_avg = p->time_steps && steps ? step_ms / (double)steps : 0.0;
        h->timing.steps = steps;
        h->timing.nnz = h->nnz;
        h->timing.n = n;
        h->timing.R = R;
        h->timing.M = M;
        h->timing.grid = G;
        h->timing.kernel = kernel_used;
        return check_moments(h, mu, R, M);
    });
}

// The same moments as leo2d_kpm_moments (FP64, reference row order) with every
// device buffer carved out of the library's static device scratch: no device
// memory is allocated or freed.  Layout of one block of B vectors, each region
// 256-byte aligned: T_j and T_{j+1} blocks [n][B], the start vectors (reference
// scheme with caller vectors), the upload staging of caller vectors [B][w n],
// the step partials [grid][B][3], the moments [B][M] and the targets [B].
int leo2d_kpm_moments_noalloc(leo2d_handle* h, const leo2d_kpm_params* p, double* mu)
{
    if (!h || !p || !mu) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        check_kpm_params(h, p);
        if (p->precision != LEO2D_PREC_FP64)
            throw ArgError("leo2d_kpm_moments_noalloc runs the FP64 recursion only (the FP32 matrix copy of the "
                           "mixed recursion is allocated on first use)");
        char* base = static_cast<char*>(leo2d::scratch_ptr(h->device));
        if (!base) throw StateError("static device scratch not initialised on this device");
        const int M = p->M, R = p->R, n = h->n;
        const int B = p->block ? p->block : std::min(64, next_pow2(R));
        const int G = p->grid > 0 ? p->grid : auto_grid(n, B);
        if (G % 8 != 0) throw ArgError("grid must be a multiple of 8");
        const bool real_vec = !h->is_complex && (p->start == LEO2D_START_TARGETS ||
                                                 (p->start == LEO2D_START_VECTORS && !p->vectors_complex));
        if (real_vec && p->start == LEO2D_START_RANDOM_PHASE) throw ArgError("random phases need complex vectors");
        const int mat_type = h->is_complex ? leo2d::MAT_CPX_F64 : leo2d::MAT_REAL_F64;
        const int vec_type = real_vec ? leo2d::VEC_REAL_F64 : leo2d::VEC_CPX_F64;
        int dev_scheme;
        if (p->scheme == LEO2D_SCHEME_DOUBLING) dev_scheme = 1;
        else dev_scheme = p->start == LEO2D_START_TARGETS ? 0 : 2;
        const bool keep_v = dev_scheme == 2;
        const int w = p->vectors_complex ? 2 : 1;

        // carve the workspace; reject what does not fit before touching the device
        const size_t vb = leo2d::vec_bytes(vec_type);
        const size_t block_bytes = vb * (size_t)n * (size_t)B;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t at = off;
            off += (bytes + 255) / 256 * 256;
            return at;
        };
        const size_t oA = take(block_bytes), oB = take(block_bytes);
        const size_t oV = keep_v ? take(block_bytes) : 0;
        const size_t oStage = p->start == LEO2D_START_VECTORS ? take(sizeof(double) * (size_t)w * n * B) : 0;
        const size_t oPart = take(sizeof(double) * 3 * (size_t)G * B);
        const size_t oMu = take(sizeof(double) * (size_t)M * B);
        const size_t oTg = take(sizeof(int32_t) * (size_t)B);
        if (off > leo2d::kScratchBytes)
            throw ArgError("leo2d_kpm_moments_noalloc needs " + std::to_string(off) + " bytes of workspace for n = " +
                           std::to_string(n) + ", block = " + std::to_string(B) + ", M = " + std::to_string(M) +
                           "; the static scratch holds " + std::to_string(leo2d::kScratchBytes) +
                           " (use leo2d_kpm_moments)");
        void* A = base + oA;
        void* Bf = base + oB;
        void* V = keep_v ? base + oV : nullptr;
        double* stage = p->start == LEO2D_START_VECTORS ? reinterpret_cast<double*>(base + oStage) : nullptr;
        double* part = reinterpret_cast<double*>(base + oPart);
        double* dmu = reinterpret_cast<double*>(base + oMu);
        int32_t* dtg = reinterpret_cast<int32_t*>(base + oTg);

        // the kernel leo2d_kpm_moments picks for the reference row order in FP64
        int variant = kpm_variant(false, false, B, n);
        if (variant == kVariantSell) variant = B == 16 || B == 32 ? 14 : 1;
        if (variant == 14 && B != 16 && B != 32 && B != 64 && B != 128) variant = 1;
        if (variant == 15) variant = 1;
        if (B < 64 && variant != 14) variant = 0;

        hipStream_t s = h->stream;
        for (int b0 = 0; b0 < R; b0 += B) {
            const int Rb = std::min(B, R - b0);
            check(leo2d::launch_zero_f64(dmu, (int64_t)M * B, s), "moments");
            if (p->start == LEO2D_START_TARGETS) {
                int32_t tg[128];
                for (int r = 0; r < B; ++r) tg[r] = p->targets[b0 + std::min(r, Rb - 1)];
                check(hipMemcpyAsync(dtg, tg, sizeof(int32_t) * B, hipMemcpyHostToDevice, s), "targets");
                check(hipStreamSynchronize(s), "targets");
                check(leo2d::launch_init_targets(A, vec_type, n, B, dtg, s), "init");
            } else if (p->start == LEO2D_START_RANDOM_PHASE) {
                check(leo2d::launch_init_random_phase_perm(A, vec_type, n, B, p->seed, b0, nullptr, s), "init");
            } else {
                const size_t per = (size_t)w * n;
                check(hipMemcpyAsync(stage, p->vectors + per * b0, sizeof(double) * per * Rb, hipMemcpyHostToDevice, s),
                      "upload");
                check(leo2d::launch_zero_f64(stage + per * Rb, (int64_t)(per * (B - Rb)), s), "upload");
                check(leo2d::launch_init_from_host_layout_perm(A, vec_type, stage, p->vectors_complex, n, B, nullptr, s),
                      "init");
            }
            if (keep_v) check(hipMemcpyAsync(V, A, block_bytes, hipMemcpyDeviceToDevice, s), "copy start");
            leo2d::StepArgs sa{};
            sa.row_ptr = (const int64_t*)h->rp.p;
            sa.col = (const int32_t*)h->col.p;
            sa.val = h->val.p;
            sa.n = n;
            sa.R = B;
            sa.part = part;
            sa.grid = G;
            sa.v = V;
            sa.variant = variant;
            sa.xcd_rows = kpm_sched() == 1 ? (n + 7) / 8 : 0;
            leo2d::ReduceArgs ra{};
            ra.part = part;
            ra.grid = G;
            ra.R = B;
            ra.M = M;
            ra.scheme = dev_scheme;
            ra.targets = dtg;
            ra.vec_type = vec_type;
            ra.mu = dmu;
            for (int j = 0; j < M / 2; ++j) {
                void* cur = (j % 2 == 0) ? A : Bf;
                void* other = (j % 2 == 0) ? Bf : A;
                sa.cur = cur;
                sa.prev = other;
                sa.out = other;
                sa.alpha = j == 0 ? 1.0 : 2.0;
                sa.beta = j == 0 ? 0 : 1;
                check(leo2d::launch_kpm_step(mat_type, vec_type, sa, s), "kpm step");
                ra.j = j;
                ra.cur = cur;
                ra.out = other;
                check(leo2d::launch_kpm_reduce(ra, s), "kpm reduce");
            }
            check(hipMemcpyAsync(mu + (size_t)M * b0, dmu, sizeof(double) * (size_t)M * Rb, hipMemcpyDeviceToHost, s),
                  "moments");
            check(hipStreamSynchronize(s), "kpm recursion");
        }
        return check_moments(h, mu, R, M);
    });
}

int leo2d_last_timing(const leo2d_handle* h, leo2d_timing* t)
{
    if (!h || !t) return LEO2D_ERR_ARG;
    *t = h->timing;
    return LEO2D_OK;
}

int leo2d_dos(leo2d_handle* h, const double* mu, int32_t T, int32_t M, double a, double b, double scale, double* E,
              double* rho)
{
    if (!h || !mu || !rho || T < 1 || M < 2) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!(a > 0.0)) throw ArgError("energy rescale must be positive");
        hipStream_t s = h->stream;
        DevBuf dm, de, dr;
        to_device(dm, mu, sizeof(double) * (size_t)T * M, s, "moments");
        dalloc(de, sizeof(double) * M, "energies");
        dalloc(dr, sizeof(double) * (size_t)T * M, "dos");
        check(leo2d::launch_dos((const double*)dm.p, T, M, a, b, scale, (double*)de.p, (double*)dr.p, s), "dos");
        check(hipMemcpyAsync(rho, dr.p, sizeof(double) * (size_t)T * M, hipMemcpyDeviceToHost, s), "dos");
        if (E) check(hipMemcpyAsync(E, de.p, sizeof(double) * M, hipMemcpyDeviceToHost, s), "dos");
        check(hipStreamSynchronize(s), "dos");
        return LEO2D_OK;
    });
}

// ---- whole deck ---------------------------------------------------------------
int leo2d_run_deck(leo2d_handle* h, const char* deck_text, const char* out_prefix, int32_t pattern)
{
    if (!h || !deck_text) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        leo2d_geom geo;
        geo.deck = leo2d::parse_deck(deck_text);
        leo2d::setup_supercell(geo.deck);
        geo.g = leo2d::build_geometry(geo.deck);
        const leo2d::Params& o = geo.deck.opts;
        if (o.geti("diagonalize") != 0) throw leo2d::DeckError("DIAGONALIZE = 1 is not on this path");
        const bool cond = o.geti("observable_type") == 1;
        if (cond && o.geti("cond_poly_par") != 0) throw leo2d::DeckError("COND_POLY_PAR = 1 is not on this path");
        if (cond && o.geti("kpm_trace") == 1) throw leo2d::DeckError("KPM_TRACE with OBSERVABLE_TYPE = COND is not on this path");
        const std::string prefix = out_prefix ? std::string(out_prefix) : o.s.at("job_name");
        std::vector<leo2d_job> jobs = leo2d::build_jobs(geo.deck, geo.g);
        int rc = leo2d_set_geometry(h, &geo);
        if (rc != LEO2D_OK) return rc;
        rc = leo2d_pairs_build(h, nullptr, nullptr, nullptr);
        if (rc != LEO2D_OK) return rc;
        if (o.geti("strain_type") == 1) {
            // STRAIN_TYPE = SUPERCELL: StrainCalc::loadFourierConfigFile on STRAIN_FILE (locality.cpp:3593-3597)
            auto sf = o.s.find("strain_file");
            if (sf == o.s.end()) throw leo2d::DeckError("STRAIN_TYPE = SUPERCELL needs STRAIN_FILE");
            int32_t mk = 0;
            int64_t cnt = 0;
            if (leo2d_read_relaxation(sf->second.c_str(), &mk, nullptr, &cnt) != LEO2D_OK)
                throw leo2d::DeckError("cannot read STRAIN_FILE " + sf->second);
            std::vector<double> coeffs((size_t)cnt);
            if (leo2d_read_relaxation(sf->second.c_str(), &mk, coeffs.data(), &cnt) != LEO2D_OK)
                throw leo2d::DeckError("STRAIN_FILE " + sf->second + " is short of coefficients");
            rc = leo2d_set_relaxation(h, mk, coeffs.data());
            if (rc != LEO2D_OK) return rc;
        }
        const int M = (int)o.geti("poly_order");
        const bool trace = o.geti("kpm_trace") == 1;
        int samples = 0;
        std::default_random_engine generator;                          // locality.cpp:2275-2276
        std::normal_distribution<double> distribution(0.0, 1.0);
        if (trace) {
            samples = (int)o.geti("kpm_trace_samps");
            if (samples == -1) {
                samples = (int)((int)1.0e6 / ((double)geo.g.n));        // locality.cpp:2257-2270
                if (samples < 1) samples = 1;
                if (samples > 100) samples = 100;
            }
        }
        std::vector<std::vector<double>> rho_all;
        for (size_t q = 0; q < jobs.size(); ++q) {
            const leo2d_job& job = jobs[q];
            int64_t nnz = 0;
            if (cond) {
                // Locality::computeCondKPM + conductivityTransform (locality.cpp:3133-3147); the
                // reference's conductivity path takes a real H only (locality.cpp:6470)
                if (job.complex_matrix) throw leo2d::DeckError("OBSERVABLE_TYPE = COND needs a real Hamiltonian");
                rc = leo2d_assemble_velocity(h, &job, pattern, &nnz);
                if (rc != LEO2D_OK) return rc;
                const int T = job.num_targets;
                std::vector<double> cm((size_t)T * M * M);
                if (T > 0) {
                    leo2d_cond_params cp{};
                    cp.M = M;
                    cp.R = T;
                    cp.targets = job.targets;
                    rc = leo2d_cond_moments(h, &cp, cm.data());
                    if (rc != LEO2D_OK) return rc;
                    if (o.geti("cond_transform") == 1) {
                        std::vector<double> tr(cm.size());
                        rc = leo2d_cond_transform(h, M, T, cm.data(), tr.data());
                        if (rc != LEO2D_OK) return rc;
                        cm.swap(tr);
                    }
                }
                rho_all.push_back(cm);
                continue;
            }
            // the last job's velocity matrices feed the MATRIX_SAVE = 2 dumps (H is the same either way)
            if (q + 1 == jobs.size() && o.geti("matrix_save") > 1) rc = leo2d_assemble_velocity(h, &job, pattern, &nnz);
            else rc = leo2d_assemble(h, &job, pattern, &nnz);
            if (rc != LEO2D_OK) return rc;
            leo2d_kpm_params kp{};
            kp.M = M;
            kp.precision = LEO2D_PREC_FP64;
            kp.scheme = LEO2D_SCHEME_REFERENCE;
            std::vector<double> vecs;
            if (!trace) {
                kp.R = job.num_targets;
                kp.start = LEO2D_START_TARGETS;
                kp.targets = job.targets;
            } else {
                vecs.resize((size_t)samples * geo.g.n);
                for (size_t t = 0; t < vecs.size(); ++t) vecs[t] = distribution(generator);
                kp.R = samples;
                kp.start = LEO2D_START_VECTORS;
                kp.vectors = vecs.data();
            }
            std::vector<double> mu((size_t)kp.R * M);
            rc = leo2d_kpm_moments(h, &kp, mu.data());
            if (rc != LEO2D_OK) return rc;
            const double a = job.energy_rescale, b = job.energy_shift;
            std::vector<double> rho;
            if (!trace) {
                if (o.geti("dos_transform") == 1) {
                    rho.resize(mu.size());
                    rc = leo2d_dos(h, mu.data(), kp.R, M, a, b, 1.0, nullptr, rho.data());
                    if (rc != LEO2D_OK) return rc;
                } else {
                    rho = mu;
                }
            } else {
                // densityTransformTrace: average the moments first (param_tools.cpp:732-745)
                std::vector<double> avg(M, 0.0);
                for (int i = 0; i < M; ++i)
                    for (int smp = 0; smp < samples; ++smp) avg[i] = avg[i] + mu[(size_t)smp * M + i] / ((double)samples);
                rho.resize(M);
                rc = leo2d_dos(h, avg.data(), 1, M, a, b, 1.0 / (double)geo.g.n, nullptr, rho.data());
                if (rc != LEO2D_OK) return rc;
                if (o.geti("verbose_save") == 1) rho = mu;   // verbose save prints the sample moments
            }
            rho_all.push_back(rho);
        }
        // files of the last job, as every job of the reference writes them in turn
        // (locality.cpp:4360-4477): _matrix.dat, and with MATRIX_POS_SAVE the positions
        if (!jobs.empty() && o.geti("matrix_save") > 0) {
            int32_t nn;
            int64_t z;
            int32_t cpx;
            leo2d_csr_info(h, &nn, &z, &cpx);
            std::vector<int64_t> rp(nn + 1);
            std::vector<int32_t> ci((size_t)z);
            std::vector<double> vv((size_t)z * (cpx ? 2 : 1));
            rc = leo2d_csr_copy(h, rp.data(), ci.data(), vv.data());
            if (rc != LEO2D_OK) return rc;
            {
                std::ofstream f(prefix + "_matrix.dat");
                f << leo2d::matrix_dat_text(nn, rp.data(), ci.data(), vv.data(), cpx);
            }
            if (o.geti("matrix_save") > 1) {
                // locality.cpp:4385-4416 (real: dxH and dyH), 5230-5244 (complex: dxH only),
                // written like _matrix.dat
                std::vector<double> vx(vv.size()), vy(vv.size());
                rc = leo2d_velocity_copy(h, vx.data(), vy.data());
                if (rc != LEO2D_OK) return rc;
                {
                    std::ofstream f(prefix + "_dxH_matrix.dat");
                    f << leo2d::matrix_dat_text(nn, rp.data(), ci.data(), vx.data(), cpx);
                }
                if (!cpx) {
                    std::ofstream f(prefix + "_dyH_matrix.dat");
                    f << leo2d::matrix_dat_text(nn, rp.data(), ci.data(), vy.data(), cpx);
                }
            }
        }
        if (!jobs.empty() && o.geti("matrix_pos_save") > 0) {
            const int nn = geo.g.n;
            std::vector<double> pos(3 * (size_t)nn);
            check(hipMemcpy(pos.data(), h->i2pos.p, sizeof(double) * pos.size(), hipMemcpyDeviceToHost), "positions");
            std::vector<int32_t> i2g(geo.g.i2g.begin(), geo.g.i2g.end());
            {
                std::ofstream f(prefix + "_pos.dat");
                f << leo2d::pos_dat_text(nn, pos.data(), i2g.data());
            }
            if (o.geti("matrix_pos_save") > 1) {
                const int64_t counts[2] = {h->n_intra, h->n_pairs - h->n_intra};
                const char* ext[2] = {"_intra_pos.dat", "_inter_pos.dat"};
                for (int which = 0; which < 2; ++which) {
                    std::vector<int32_t> ij(2 * (size_t)std::max<int64_t>(counts[which], 1));
                    rc = leo2d_pairs_copy(h, which, ij.data(), nullptr, nullptr);
                    if (rc != LEO2D_OK) return rc;
                    std::ofstream f(prefix + ext[which]);
                    f << leo2d::pair_pos_text(counts[which], ij.data(), pos.data());
                }
            }
        }
        {
            std::ofstream f(prefix + ".cheb");
            f << leo2d::cheb_text(geo.deck, jobs, rho_all, M);
        }
        // the job list's Bloch vectors (locality.cpp:2785-2797)
        if (o.geti("k_sampling") == 1) {
            std::ofstream f(prefix + ".kpts");
            f << leo2d::kpts_text(jobs);
        }
        return LEO2D_OK;
    });
}

// ---- batched jobs ------------------------------------------------------------
int leo2d_batch_clear(leo2d_handle* h)
{
    if (!h) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        check(hipStreamSynchronize(h->stream), "batch clear");
        h->batch.drop_graph();
        h->batch = leo2d_handle::Batch();
        return LEO2D_OK;
    });
}

int leo2d_batch_append(leo2d_handle* h, int32_t* job)
{
    if (!h) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_csr) throw StateError("no Hamiltonian on the device");
        auto& B = h->batch;
        if (B.is_complex >= 0 && B.is_complex != h->is_complex)
            throw ArgError("a batch holds either real or complex Hamiltonians, not both");
        leo2d_handle::Batch::Pending q;
        q.n = h->n;
        q.nnz = h->nnz;
        const size_t vb = h->is_complex ? 16 : 8;
        hipStream_t s = h->stream;
        dalloc(q.rp, sizeof(int64_t) * (size_t)(q.n + 1), "batch job rows");
        dalloc(q.col, sizeof(int32_t) * (size_t)q.nnz, "batch job columns");
        dalloc(q.val, vb * (size_t)q.nnz, "batch job values");
        check(hipMemcpyAsync(q.rp.p, h->rp.p, sizeof(int64_t) * (size_t)(q.n + 1), hipMemcpyDeviceToDevice, s), "append");
        if (q.nnz) {
            check(hipMemcpyAsync(q.col.p, h->col.p, sizeof(int32_t) * (size_t)q.nnz, hipMemcpyDeviceToDevice, s), "append");
            check(hipMemcpyAsync(q.val.p, h->val.p, vb * (size_t)q.nnz, hipMemcpyDeviceToDevice, s), "append");
        }
        check(hipStreamSynchronize(s), "append");
        B.is_complex = h->is_complex;
        B.pending.push_back(std::move(q));
        if (job) *job = B.jobs() - 1;
        return LEO2D_OK;
    });
}

int leo2d_batch_info(const leo2d_handle* h, int32_t* jobs, int64_t* rows, int64_t* nnz)
{
    if (!h) return LEO2D_ERR_ARG;
    const auto& B = h->batch;
    int64_t r = B.rows, z = B.nnz;
    for (const auto& q : B.pending) {
        r += q.n;
        z += q.nnz;
    }
    if (jobs) *jobs = B.jobs();
    if (rows) *rows = r;
    if (nnz) *nnz = z;
    return LEO2D_OK;
}

int leo2d_batch_kpm(leo2d_handle* h, const leo2d_batch_params* p, double* mu)
{
    if (!h || !p || !mu) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        auto& B = h->batch;
        if (B.jobs() == 0) throw StateError("the batch holds no jobs");
        const int M = p->M, R = p->R;
        if (M < 4 || M % 4 != 0) throw ArgError("M must be a positive multiple of 4");
        if (R != 4 && R != 8 && R != 16 && R != 32) throw ArgError("R must be 4, 8, 16 or 32");
        if (p->precision != LEO2D_PREC_FP64 && p->precision != LEO2D_PREC_MIXED) throw ArgError("unknown precision");
        if (p->mode != LEO2D_BATCH_ROW_GROUPS && p->mode != LEO2D_BATCH_ENTRY_SLOTS && p->mode != LEO2D_BATCH_ROW_CHUNKS &&
            p->mode != LEO2D_BATCH_SELL)
            throw ArgError("unknown mode");
        if (p->mode == LEO2D_BATCH_ROW_CHUNKS && R != 16 && R != 32) throw ArgError("row chunks need R = 16 or 32");
        constexpr int kMinTile = 16;
        if (p->tile_rows != 0 && p->tile_rows < kMinTile) throw ArgError("tile_rows must be 0 or at least 16");
        if (p->grid < 0 || p->grid % 8 != 0) throw ArgError("grid must be 0 or a positive multiple of 8");
        if (p->start != LEO2D_START_RANDOM_PHASE && p->start != LEO2D_START_TARGETS) throw ArgError("unknown start");
        if (p->start == LEO2D_START_TARGETS && !p->targets) throw ArgError("LEO2D_START_TARGETS needs targets");
        batch_stack(h);
        hipStream_t s = h->stream;
        const int P = (int)B.job_row0.size() - 1;
        const bool mixed = p->precision == LEO2D_PREC_MIXED;
        const int32_t* dtargets = nullptr;
        if (p->start == LEO2D_START_TARGETS) {
            for (int q = 0; q < P; ++q)
                for (int r = 0; r < R; ++r) {
                    const int32_t t = p->targets[(size_t)q * R + r];
                    if (t < 0 || t >= B.job_row0[q + 1] - B.job_row0[q])
                        throw ArgError("target " + std::to_string(t) + " outside job " + std::to_string(q));
                }
            if (B.tgt.bytes != std::max<size_t>(sizeof(int32_t) * (size_t)P * R, 16)) {
                B.drop_graph();
                dalloc(B.tgt, sizeof(int32_t) * (size_t)P * R, "batch targets");
            }
            check(hipMemcpyAsync(B.tgt.p, p->targets, sizeof(int32_t) * (size_t)P * R, hipMemcpyHostToDevice, s),
                  "batch targets");
            dtargets = (const int32_t*)B.tgt.p;
        }
        const int mat_type = B.is_complex ? (mixed ? leo2d::MAT_CPX_F32 : leo2d::MAT_CPX_F64)
                                          : (mixed ? leo2d::MAT_REAL_F32 : leo2d::MAT_REAL_F64);
        // a real H from unit-vector starts keeps real vectors (as the single-job path does)
        const bool real_vec = !B.is_complex && p->start == LEO2D_START_TARGETS;
        const int vec_type = real_vec ? (mixed ? leo2d::VEC_REAL_F32 : leo2d::VEC_REAL_F64)
                                      : (mixed ? leo2d::VEC_CPX_F32 : leo2d::VEC_CPX_F64);
        const void* dval = B.val.p;
        if (mixed) {
            if (!B.val32_valid) {
                const size_t count = (size_t)B.nnz * (B.is_complex ? 2 : 1);
                dalloc(B.val32, sizeof(float) * count, "batch fp32 values");
                check(leo2d::launch_convert_f64_to_f32((const double*)B.val.p, (float*)B.val32.p, (int64_t)count, s),
                      "batch fp32 values");
                B.val32_valid = true;
                B.drop_graph();
            }
            dval = B.val32.p;
        }
        // device buffers, kept between calls so a captured graph stays valid
        const size_t vbytes = leo2d::vec_bytes(vec_type) * (size_t)B.rows * R;
        auto ensure = [&](DevBuf& b, size_t bytes, const char* what) {
            if (b.bytes != std::max<size_t>(bytes, 16)) {
                B.drop_graph();
                dalloc(b, bytes, what);
            }
        };
        // SELL tiles of the stack (complex vectors, R = 16 or 32; FP32 only at 32): tiles of at most
        // 56 rows for the kernel with a copy wavefront, else 64 rows whose union fits the LDS the
        // eight-wavefront kernel leaves beside its T_{j-1} copies and partials
        // (built on the host once per stack, so only when mode 3 is asked for or may be tuned in)
        const bool sell_cw = env_int("LEO2D_SELL_CW", 1) != 0;
        const bool want_sell = p->mode == LEO2D_BATCH_SELL || p->tune;
        // the copy wavefront addresses a vector chunk with 32-bit byte offsets
        const bool sell_fits = !sell_cw || (uint64_t)B.rows * (uint64_t)R * leo2d::vec_bytes(vec_type) < (uint64_t(1) << 32);
        const bool sell_ok = want_sell && leo2d::sell_supported(mat_type, vec_type, R, sell_cw) && R <= 32 && sell_fits &&
                             build_sell_tiles(B.sell, h->sell_stage, s, (int)B.rows, B.nnz, (const int64_t*)B.rp.p,
                                              (const int32_t*)B.col.p,
                                              sell_cw ? leo2d::kSellUcap : leo2d::kSellUcapTiled,
                                              sell_cw ? leo2d::kSellCwRows : leo2d::kSellMaxRows, false, &B.job_row0);
        if (!p->tune && p->mode == LEO2D_BATCH_SELL && !sell_ok)
            throw ArgError("LEO2D_BATCH_SELL needs complex vectors with R = 16 or 32 (FP32 at 16 only with the copy "
                           "wavefront) or 32 real FP64 vectors, rows of at most 64 entries, and with the copy wavefront "
                           "a stack of rows x R x vector bytes below 4 GiB");
        int64_t max_tiles = 0;   // tiles of the shortest allowed height bound every tiling's count
        for (int q = 0; q < P; ++q) max_tiles += (B.job_row0[q + 1] - B.job_row0[q] + kMinTile - 1) / kMinTile;
        if (sell_ok) max_tiles = std::max<int64_t>(max_tiles, B.sell.ntiles);
        ensure(B.A, vbytes, "batch vectors");
        ensure(B.B, vbytes, "batch vectors");
        ensure(B.part, sizeof(double) * 2 * (size_t)std::max<int64_t>(max_tiles, 1) * R, "batch partials");
        ensure(B.mu, sizeof(double) * (size_t)P * R * M, "batch moments");

        leo2d::BatchStepArgs sa{};
        sa.row_ptr = (const int64_t*)B.rp.p;
        sa.col = (const int32_t*)B.col.p;
        sa.val = dval;
        sa.R = R;
        sa.part = (double*)B.part.p;
        // the SELL step: one workgroup per CU, partials per tile
        leo2d::SellStepArgs ss{};
        int cus = 256;
        {
            int dev_id = 0;
            if (hipGetDevice(&dev_id) == hipSuccess)
                (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id);
        }
        if (sell_ok) {
            auto& S = B.sell;
            const int hb = mat_type == leo2d::MAT_CPX_F64 ? 16 : (mat_type == leo2d::MAT_REAL_F32 ? 4 : 8);
            ss.ev = sell_values(S, s, dval, hb, mixed);
            ss.n = (int)B.rows;
            ss.R = R;
            ss.part = (double*)B.part.p;
            ss.grid = std::max(8, cus / 8 * 8);
            ss.tile_part = 1;
            ss.copy_wave = sell_cw ? 1 : 0;
            ss.ucap = S.ucap;
            ss.ntiles = S.ntiles;
            ss.tile_row0 = (const int32_t*)S.row0.p;
            ss.tile_u0 = (const int64_t*)S.u0.p;
            ss.ucols = (const int32_t*)S.ucols.p;
            ss.tile_g0 = (const int32_t*)S.g0.p;
            ss.grow = (const int32_t*)S.grow.p;
            ss.glen = (const int32_t*)S.glen.p;
            ss.goff = (const int64_t*)S.goff.p;
            ss.eu = (const uint16_t*)S.eu.p;
            ss.tile_rows = S.tile_rows;
            ss.max_len = S.max_len;
        }
        const int32_t* job_tiles = (const int32_t*)B.djob_tile0.p;
        auto set_shape = [&](int mode, int T, int grid) {
            sa.mode = mode;
            if (mode == LEO2D_BATCH_SELL) {
                sa.grid = ss.grid;
                job_tiles = (const int32_t*)B.sell.djob_tile0.p;
                return;
            }
            batch_tiles(h, T);
            sa.tile_row0 = (const int32_t*)B.dtile_row0.p;
            sa.ntiles = B.ntiles;
            sa.grid = grid > 0 ? grid : batch_default_grid(B.ntiles);
            job_tiles = (const int32_t*)B.djob_tile0.p;
        };
        auto step = [&](void* cur, void* out, int j) {
            if (sa.mode == LEO2D_BATCH_SELL) {
                ss.cur = cur;
                ss.prev = out;
                ss.out = out;
                ss.alpha = j == 0 ? 1.0 : 2.0;
                ss.beta = j == 0 ? 0 : 1;
                ss.reverse = j & 1;
                check(leo2d::launch_kpm_step_sell(mat_type, vec_type, ss, s), "batch step");
                return;
            }
            sa.cur = cur;
            sa.out = out;
            sa.alpha = j == 0 ? 1.0 : 2.0;
            sa.beta = j == 0 ? 0 : 1;
            check(leo2d::launch_batch_step(mat_type, vec_type, sa, s), "batch step");
        };
        hipEvent_t e0 = event_at(h, 0), e1 = event_at(h, 1);

        // launch shape: given, cached from an earlier tuning, or tuned now
        int mode = p->mode, T = p->tile_rows > 0 ? p->tile_rows : 128, grid = p->grid;
        double tune_ms = 0.0;
        int candidates = 0;
        if (p->tune) {
            const leo2d_handle::Batch::Tuned* hit = nullptr;
            for (const auto& t : B.tuned)
                if (t.R == R && t.precision == p->precision && t.vec_type == vec_type) hit = &t;
            if (hit) {
                mode = hit->mode;
                T = hit->tile_rows;
                grid = hit->grid;
            } else {
                // time a few Chebyshev steps for every candidate shape on this batch
                check(leo2d::launch_batch_init(B.A.p, vec_type, (const int64_t*)B.drow0.p, P, B.rows, R, p->seed, dtargets, s),
                      "batch init");
                set_shape(0, 128, 0);
                step(B.A.p, B.B.p, 0);
                check(hipEventRecord(e0, s), "event");
                float best = 1e30f;
                if (sell_ok) {
                    // SELL tiles: the shape is fixed by the LDS budget (one candidate)
                    set_shape(LEO2D_BATCH_SELL, 0, 0);
                    step(B.A.p, B.B.p, 1);
                    check(hipEventRecord(e1, s), "event");
                    for (int k = 0; k < 4; ++k) step(k % 2 ? B.B.p : B.A.p, k % 2 ? B.A.p : B.B.p, 1 + k);
                    hipEvent_t e2 = event_at(h, 2);
                    check(hipEventRecord(e2, s), "event");
                    check(hipEventSynchronize(e2), "tune");
                    float ms = 0.f;
                    check(hipEventElapsedTime(&ms, e1, e2), "event");
                    ++candidates;
                    best = ms;
                    mode = LEO2D_BATCH_SELL;
                    T = B.sell.tile_rows;
                    grid = ss.grid;
                }
                for (int tt : {64, 128, 256, 512}) {
                    for (int md : {0, 1, 2}) {
                        if (md == 2 && R != 16 && R != 32) continue;
                        int last = -1;
                        for (int per_cu : {2, 4, 8}) {
                            const int g = std::min(batch_default_grid((int)((B.rows + tt - 1) / tt) + P), 256 * per_cu);
                            if (g == last) continue;
                            last = g;
                            set_shape(md, tt, g);
                            step(B.A.p, B.B.p, 1);
                            check(hipEventRecord(e1, s), "event");
                            for (int k = 0; k < 4; ++k) step(k % 2 ? B.B.p : B.A.p, k % 2 ? B.A.p : B.B.p, 1);
                            hipEvent_t e2 = event_at(h, 2);
                            check(hipEventRecord(e2, s), "event");
                            check(hipEventSynchronize(e2), "tune");
                            float ms = 0.f;
                            check(hipEventElapsedTime(&ms, e1, e2), "event");
                            ++candidates;
                            if (ms < best) {
                                best = ms;
                                mode = md;
                                T = tt;
                                grid = sa.grid;
                            }
                        }
                    }
                }
                hipEvent_t e3 = event_at(h, 3);
                check(hipEventRecord(e3, s), "event");
                check(hipEventSynchronize(e3), "tune");
                float tms = 0.f;
                check(hipEventElapsedTime(&tms, e0, e3), "event");
                tune_ms = tms;
                B.tuned.push_back({R, p->precision, vec_type, mode, T, grid});
            }
        }
        set_shape(mode, T, grid);
        grid = sa.grid;
        if (mode == LEO2D_BATCH_SELL) T = B.sell.tile_rows;

        auto record = [&]() {
            check(leo2d::launch_batch_init(B.A.p, vec_type, (const int64_t*)B.drow0.p, P, B.rows, R, p->seed, dtargets, s),
                  "batch init");
            for (int j = 0; j < M / 2; ++j) {
                void* cur = (j % 2 == 0) ? B.A.p : B.B.p;
                void* other = (j % 2 == 0) ? B.B.p : B.A.p;
                step(cur, other, j);
                check(leo2d::launch_batch_reduce((const double*)B.part.p, job_tiles, P, R, M, j, (double*)B.mu.p, s),
                      "batch reduce");
            }
        };
        if (p->graph) {
            const leo2d_handle::Batch::GraphKey key{M, R, p->precision, mode, T, grid, p->seed,
                                                     B.A.p, B.B.p, B.part.p, B.mu.p,
                                                     mode == LEO2D_BATCH_SELL ? ss.ev : dval, dtargets,
                                                     mode == LEO2D_BATCH_SELL ? ss.copy_wave : 0};
            if (!B.exec || !(key == B.key)) {
                B.drop_graph();
                hipGraph_t graph = nullptr;
                check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "graph capture");
                try {
                    record();
                } catch (...) {
                    (void)hipStreamEndCapture(s, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                check(hipStreamEndCapture(s, &graph), "graph capture");
                const hipError_t ie = hipGraphInstantiate(&B.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) {
                    B.exec = nullptr;
                    check(ie, "graph instantiate");
                }
                B.key = key;
            }
            check(hipEventRecord(e0, s), "event");
            check(hipGraphLaunch(B.exec, s), "graph launch");
            check(hipEventRecord(e1, s), "event");
        } else {
            check(hipEventRecord(e0, s), "event");
            record();
            check(hipEventRecord(e1, s), "event");
        }
        check(hipStreamSynchronize(s), "batch recursion");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, e0, e1), "event");
        check(hipMemcpy(mu, B.mu.p, sizeof(double) * (size_t)P * R * M, hipMemcpyDeviceToHost), "batch moments");

        auto& t = B.timing;
        t.recursion_ms = ms;
        t.tune_ms = tune_ms;
        t.mode = mode;
        t.tile_rows = T;
        t.grid = grid;
        t.graph = p->graph ? 1 : 0;
        t.candidates = candidates;
        t.jobs = P;
        t.ntiles = mode == LEO2D_BATCH_SELL ? B.sell.ntiles : B.ntiles;
        t.rows = B.rows;
        t.nnz = B.nnz;
        for (int q = 0; q < P * R; ++q) {
            const double* m = mu + (size_t)q * M;
            const double lim = std::fabs(m[0]) * (1.0 + 1e-6) + 1e-12;
            for (int j = 0; j < M; ++j) {
                if (!std::isfinite(m[j])) {
                    h->err = "non-finite moment (job " + std::to_string(q / R) + ", vector " + std::to_string(q % R) + ")";
                    return LEO2D_ERR_NAN;
                }
                if (std::fabs(m[j]) > lim) {
                    h->err = "job " + std::to_string(q / R) + ": |mu_" + std::to_string(j) +
                             "| > mu_0: rescaled spectrum leaves [-1, 1]";
                    return LEO2D_ERR_SPECTRUM;
                }
            }
        }
        return LEO2D_OK;
    });
}

int leo2d_batch_last_timing(const leo2d_handle* h, leo2d_batch_timing* t)
{
    if (!h || !t) return LEO2D_ERR_ARG;
    *t = h->batch.timing;
    return LEO2D_OK;
}



int leo2d_cond_moments(leo2d_handle* h, const leo2d_cond_params* p, double* C)
{
    if (!h || !p || !C) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        if (!h->has_csr || !h->has_dx) throw StateError("leo2d_cond_moments needs leo2d_assemble_velocity");
        if (h->is_complex) throw ArgError("the conductivity recursion takes a real H");
        const int M = p->M, R = p->R, n = h->n;
        if (M < 2) throw ArgError("M must be at least 2");
        if (R <= 0 || !p->targets) throw ArgError("conductivity needs R >= 1 targets");
        for (int r = 0; r < R; ++r)
            if (p->targets[r] < 0 || p->targets[r] >= n) throw ArgError("target outside the matrix");
        // targets per device block: a width the step kernels take, bounded so the two
        // vector sets (2 x B x M x n doubles) stay within 96 GiB
        int want = p->block > 0 ? p->block : std::min(R, 16);
        int B = 1;
        while (B * 2 <= std::min(want, 128)) B *= 2;
        const double budget = 96.0 * (1ull << 30);
        while (B > 1 && 2.0 * B * (double)M * n * sizeof(double) > budget) B /= 2;
        if (2.0 * B * (double)M * n * sizeof(double) > budget) throw ArgError("M x n too large for one target's sets");
        hipStream_t s = h->stream;
        const int64_t* drp = (const int64_t*)h->rp.p;
        const int32_t* dcol = (const int32_t*)h->col.p;
        const size_t nb = (size_t)n * B;
        DevBuf Bset, Dset, X0, X1, dtg, Cdev;
        dalloc(Bset, sizeof(double) * nb * M, "beta set");
        dalloc(Dset, sizeof(double) * nb * M, "alpha set");
        dalloc(X0, sizeof(double) * nb, "vector block");
        dalloc(X1, sizeof(double) * nb, "vector block");
        dalloc(dtg, sizeof(int32_t) * B, "targets");
        dalloc(Cdev, sizeof(double) * (size_t)B * M * M, "conductivity moments");
        // split the n-long GEMM reduction into S row chunks so that the M x M outputs of a pass
        // fill every CU twice over with 128 x 128 tiles; partial products summed in chunk order
        int dev_id = 0, cus = 256;
        if (hipGetDevice(&dev_id) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id);
        const int tiles = std::max(1, ((M + 127) / 128) * ((M + 127) / 128) * B);
        const int S = std::max(1, std::min({64, (2 * cus + tiles - 1) / tiles, n / 1024}));
        DevBuf Cpart;
        if (S > 1) dalloc(Cpart, sizeof(double) * (size_t)B * S * M * M, "conductivity partials");
        auto scatter = [&](const void* blk, int m, DevBuf& set) {
            check(leo2d::launch_cond_scatter((const double*)blk, n, B, m, M, (double*)set.p, s), "conductivity scatter");
        };
        const void* dH = h->val.p;
        const void* dX = h->val_dx.p;
        hipEvent_t e0 = event_at(h, 0), e1 = event_at(h, 1), e2 = event_at(h, 2);
        float rec_ms = 0.f, gemm_ms = 0.f;
        for (int b0 = 0; b0 < R; b0 += B) {
            const int Bb = std::min(B, R - b0);
            std::vector<int32_t> tg(B);
            for (int b = 0; b < B; ++b) tg[b] = p->targets[b0 + std::min(b, Bb - 1)];
            check(hipMemcpyAsync(dtg.p, tg.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, s), "targets");
            check(hipEventRecord(e0, s), "event");
            // beta_m = T_m(H) e_t, written into the set by the step itself
            check(leo2d::launch_init_targets(X0.p, leo2d::VEC_REAL_F64, n, B, (const int32_t*)dtg.p, s), "init");
            scatter(X0.p, 0, Bset);
            auto dual = [&](const void* hv, const void* dv, const void* x, void* prev, void* out, double alpha, int beta,
                            int j, DevBuf* hset, DevBuf* dset) {
                check(leo2d::launch_cond_dual_step(drp, dcol, (const double*)hv, (const double*)dv, (const double*)x,
                                                   (const double*)prev, (double*)out, alpha, beta, n, B, j, M,
                                                   hset ? (double*)hset->p : nullptr, dset ? (double*)dset->p : nullptr,
                                                   s),
                      "conductivity step");
            };
            dual(dH, nullptr, X0.p, nullptr, X1.p, 1.0, 0, 1, &Bset, nullptr);
            void *cur = X1.p, *prv = X0.p;
            for (int m = 2; m < M; ++m) {
                dual(dH, nullptr, cur, prv, prv, 2.0, 1, m, &Bset, nullptr);
                std::swap(cur, prv);
            }
            // alpha_0 = -dxH e_t; a_1 = H a_0, a_j = 2 H a_{j-1} - a_{j-2}.  One pass over a_{j-1}
            // forms both a_j and D_{j-1} = dxH a_{j-1} (H and dxH share the pattern)
            check(leo2d::launch_init_targets(X0.p, leo2d::VEC_REAL_F64, n, B, (const int32_t*)dtg.p, s), "init");
            dual(dX, nullptr, X0.p, nullptr, X1.p, -1.0, 0, 0, nullptr, nullptr);
            cur = X1.p;      // a_0
            prv = X0.p;
            for (int j = 1; j < M; ++j) {
                dual(dH, dX, cur, j == 1 ? nullptr : prv, prv, j == 1 ? 1.0 : 2.0, j == 1 ? 0 : 1, j - 1, nullptr, &Dset);
                std::swap(cur, prv);
            }
            dual(nullptr, dX, cur, nullptr, nullptr, 1.0, 0, M - 1, nullptr, &Dset);
            check(hipEventRecord(e1, s), "event");
            // C_b[j][m] = sum_i D_b[j][i] Bset_b[m][i], stored column-major (m + j M): the GEMM's
            // rows are j (A = D_b, k = i contiguous), its columns m (B = Bset_b), so a wavefront's
            // 16 output columns are 16 consecutive doubles
            const size_t mm = (size_t)M * M;
            leo2d::DgemmArgs ga{};
            ga.M = M;
            ga.N = M;
            ga.K = n;
            ga.batch = Bb;
            ga.splits = S;
            ga.A = (const double*)Dset.p;
            ga.sam = n;
            ga.sak = 1;
            ga.sab = (int64_t)n * M;
            ga.B = (const double*)Bset.p;
            ga.sbn = n;
            ga.sbk = 1;
            ga.sbb = (int64_t)n * M;
            ga.C = S == 1 ? (double*)Cdev.p : (double*)Cpart.p;
            ga.scm = M;
            ga.scn = 1;
            ga.scs = (int64_t)mm;
            ga.scb = (int64_t)(S * mm);
            check(leo2d::launch_dgemm(ga, s), "conductivity GEMM");
            if (S > 1) check(leo2d::launch_cond_chunk_sum((const double*)Cpart.p, S, mm, Bb, (double*)Cdev.p, s), "chunk sum");
            check(hipEventRecord(e2, s), "event");
            check(hipMemcpyAsync(C + (size_t)b0 * M * M, Cdev.p, sizeof(double) * (size_t)Bb * M * M,
                                 hipMemcpyDeviceToHost, s),
                  "moments");
            check(hipStreamSynchronize(s), "conductivity");
            float a = 0.f, g = 0.f;
            check(hipEventElapsedTime(&a, e0, e1), "event");
            check(hipEventElapsedTime(&g, e1, e2), "event");
            rec_ms += a;
            gemm_ms += g;
        }
        for (size_t q = 0; q < (size_t)R * M * M; ++q)
            if (!std::isfinite(C[q])) throw NanError("non-finite conductivity moment");
        h->cond_timing.recursion_ms = rec_ms;
        h->cond_timing.gemm_ms = gemm_ms;
        h->cond_timing.block = B;
        return LEO2D_OK;
    });
}

int leo2d_cond_transform(leo2d_handle* h, int32_t M, int32_t T, const double* C, double* out)
{
    if (!h || !C || !out || M < 2 || T <= 0) return LEO2D_ERR_ARG;
    return guarded(h, [&]() {
        hipStream_t s = h->stream;
        const size_t mm = (size_t)M * M;
        DevBuf Cd, X, P, Yb;
        dalloc(Cd, sizeof(double) * mm * T, "conductivity moments");
        dalloc(X, sizeof(double) * mm * T, "weighted moments");
        dalloc(P, sizeof(double) * mm, "DCT-III matrix");
        dalloc(Yb, sizeof(double) * mm * T, "half transform");
        check(hipMemcpyAsync(Cd.p, C, sizeof(double) * mm * T, hipMemcpyHostToDevice, s), "upload");
        check(leo2d::launch_cond_weight((const double*)Cd.p, M, T, (double*)X.p, s), "weights");
        check(leo2d::launch_dct3_matrix(M, (double*)P.p, s), "DCT-III matrix");
        // with the row-major X read as column-major X^T, Z = P X^T P^T (column-major)
        // is the row-major 2D transform P X P^T: Y_t = P X_t, then X_t = Y_t P^T
        leo2d::DgemmArgs ga{};
        ga.M = M;
        ga.N = M;
        ga.K = M;
        ga.batch = T;
        ga.splits = 1;
        ga.A = (const double*)P.p;   // P(m, k) = P[m + k M]
        ga.sam = 1;
        ga.sak = M;
        ga.sab = 0;
        ga.B = (const double*)X.p;   // X_t(k, n) = X[t mm + k + n M]
        ga.sbn = M;
        ga.sbk = 1;
        ga.sbb = (int64_t)mm;
        ga.C = (double*)Yb.p;
        ga.scm = 1;
        ga.scn = M;
        ga.scb = (int64_t)mm;
        check(leo2d::launch_dgemm(ga, s), "transform GEMM");
        ga.A = (const double*)Yb.p;  // Y_t(m, k) = Yb[t mm + m + k M]
        ga.sab = (int64_t)mm;
        ga.B = (const double*)P.p;   // P^T(k, n) = P[n + k M]
        ga.sbn = 1;
        ga.sbk = M;
        ga.sbb = 0;
        ga.C = (double*)X.p;
        check(leo2d::launch_dgemm(ga, s), "transform GEMM");
        check(hipMemcpyAsync(out, X.p, sizeof(double) * mm * T, hipMemcpyDeviceToHost, s), "download");
        check(hipStreamSynchronize(s), "transform");
        return LEO2D_OK;
    });
}

int leo2d_cond_last_timing(leo2d_handle* h, leo2d_cond_timing* t)
{
    if (!h || !t) return LEO2D_ERR_ARG;
    t->recursion_ms = h->cond_timing.recursion_ms;
    t->gemm_ms = h->cond_timing.gemm_ms;
    t->block = h->cond_timing.block;
    t->reserved = 0;
    return LEO2D_OK;
}

}  // extern "C"
