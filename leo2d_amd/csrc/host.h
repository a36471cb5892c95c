// host.h -- host-side model of the KPM path: parsed deck (Job_params bag +
// per-sheet Sdata), geometry (sheet grids, global index, positions) and the
// job list.  Internal to libleo2d; the public boundary is include/leo2d.h.
#pragma once

#include <array>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/leo2d.h"
#include "model.h"

namespace leo2d {

struct DeckError : std::runtime_error {
    explicit DeckError(const std::string& m) : std::runtime_error(m) {}
};


// Sdata (src/geom/sdata.h): per-sheet shape snapshot taken at the MATERIAL line
struct SheetData {
    int mat = MAT_GRAPHENE;
    int min_shape[2] = {0, 0};
    int max_shape[2] = {0, 0};
    double max_R = 0.0;
    int boundary_condition = 0;
    double supercell[2][2] = {{0, 0}, {0, 0}};  // sheet-local frame
    int stride[2][2] = {{0, 0}, {0, 0}};
};

// Job_params (src/params/job_params.h): typed bag with the reference defaults
struct Params {
    std::map<std::string, long long> i;
    std::map<std::string, double> d;
    std::map<std::string, std::string> s;
    std::map<std::string, std::vector<int>> iv;
    std::map<std::string, std::vector<std::vector<double>>> dm;
    Params();
    long long geti(const std::string& k) const;
    double getd(const std::string& k) const;
    bool has(const std::string& k) const { return i.count(k) || d.count(k) || s.count(k) || iv.count(k) || dm.count(k); }
};

struct Deck {
    Params opts;
    std::vector<SheetData> sheets;
    std::vector<double> heights, angles;
    std::vector<int> has_material;
};

Deck parse_deck(const std::string& text);
void setup_supercell(Deck& d);  // Locality::setupSupercell, types 1 and 2

struct Geom {
    int num_sheets = 0;
    int norb = 2;
    int bc = 0;
    int sc_search = 1;
    double gsc[2][2] = {{0, 0}, {0, 0}};   // global supercell
    double lat[2][2];
    double apos[2][3];
    double ainv[2][2];
    std::vector<SheetData> sheets;
    std::vector<double> angle, height;
    std::vector<int> gs_on;
    std::vector<std::array<double, 2>> gshift;
    std::vector<int> H, W, count, base;
    std::vector<double> scinv;              // [s][2][2]
    std::vector<std::vector<int>> grid;     // [s][H*W*norb]
    int n = 0;
    std::vector<int> i2g;                   // [4n]
    std::vector<double> i2p;                // [3n] global frame
    std::vector<double> lpos;               // [3n] sheet-local frame
};

Geom build_geometry(const Deck& d);
// Locality::getReciprocal(a) (locality.cpp:7215-7262): b_i . a_j = 2 pi delta_ij
void reciprocal_vectors(const double a[2][2], double b[2][2]);
std::vector<leo2d_job> build_jobs(const Deck& d, const Geom& g);
std::vector<int> target_list(const Deck& d, const Geom& g);

// writers
std::string cheb_text(const Deck& d, const std::vector<leo2d_job>& jobs, const std::vector<std::vector<double>>& rho,
                      int M);
std::string matrix_dat_text(int n, const int64_t* rp, const int32_t* col, const double* val, int is_complex);
// .kpts: one line per k-sampled job with its Bloch vector
std::string kpts_text(const std::vector<leo2d_job>& jobs);
// MATRIX_POS_SAVE writers: _pos.dat (job positions, sheet, orbital) and the pair-position lists
std::string pos_dat_text(int n, const double* i2pos, const int32_t* i2g);
std::string pair_pos_text(int64_t np, const int32_t* ij, const double* i2pos);

}  // namespace leo2d
