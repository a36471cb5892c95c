// assemble.h -- device-side geometry/job descriptors and launchers of
// assemble.hip.  Internal to libleo2d.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/leo2d.h"
#include "model.h"

namespace leo2d {


struct SheetDev {
    const int32_t* grid;   // [H][W][norb], -1 = empty
    int H, W, base, count;
    int min0, min1, max0, max1;
    int mat, bc;
    double sc[2][2];       // sheet-local supercell
    int st[2][2];          // supercell stride in unit cells
    double angle;
    double cneg, sneg;     // cos(-angle), sin(-angle) from the host libm
    int win;               // intralayer search window (cells each way)
};

struct GeomDev {
    int n, num_sheets, norb, bc, sc_search;
    int inter_win;         // interlayer search window, from sheet 1's material
    double gsc[2][2];
    double ainv[2][2];
    double lat[2][2];
    const int32_t* i2g;
    const double* i2p;
    const double* lpos;
    SheetDev s[LEO2D_MAX_SHEETS];
};

struct JobDev {
    double shift[LEO2D_MAX_SHEETS][2];
    double s1a[LEO2D_MAX_SHEETS][4];   // rotated lattice per sheet (host libm cos/sin)
    double sc[2][2];                   // job supercell (uniformly strained if requested)
    double energy_rescale, energy_shift;
    int elec_on;
    double E, B;
    double k[2];
    int complex_matrix;
    int uniform_strain;
    double u[2][2];
    int sheet_strain_on;
    double ss[LEO2D_MAX_SHEETS][2][2];   // per-sheet strain tensors
    // STRAIN_TYPE = SUPERCELL: Fourier relaxation field (coefficients [3][nc] on the device)
    int strain_type;
    const double* relax;
    int relax_max_k;
    double rb[2][2];                     // reciprocal vectors of the supercell
    double rot_cs[3][2];                 // cos, sin of r pi / 3, r = 0, 1, 2 (host libm)
    double basis_cs[2];                  // cos, sin of the angle of b1 (host libm)
};

// pair counts per row (one wavefront per row) and their exclusive scan
hipError_t pairs_count(const GeomDev& g, int32_t* cnt_intra, int64_t* cnt_total, int64_t* pair_rp, hipStream_t s);
// the pair lists, one thread per row; tcol/tsc: scratch of the same size as pcol/psc,
// holding the pairs as found
hipError_t pairs_write(const GeomDev& g, const int64_t* pair_rp, const int32_t* cnt_intra, int32_t* tcol, int32_t* tsc,
                       int32_t* pcol, int32_t* psc, int64_t* nslots, int64_t* slot_rp, hipStream_t s);
// the same lists, one wavefront per row with the lists staged in LDS (no scratch); only
// when pairs_staged_fit holds for the longest intralayer and interlayer lists
bool pairs_staged_fit(const GeomDev& g, int max_intra, int max_inter);
hipError_t pairs_write_staged(const GeomDev& g, const int64_t* pair_rp, int32_t* pcol, int32_t* psc, int64_t* nslots,
                              int64_t* slot_rp, hipStream_t s);
hipError_t job_positions(const GeomDev& g, const JobDev& j, double* i2pos, hipStream_t s);
// each orbital's local twist from its nearest neighbour on positions i2pos
hipError_t local_theta(const GeomDev& g, const double* i2pos, double* theta, hipStream_t s);
hipError_t fill_slots(const GeomDev& g, const JobDev& j, const double* i2pos, const int64_t* pair_rp,
                      const int32_t* cnt_intra, const int32_t* pcol, const int32_t* psc, const int64_t* slot_rp,
                      int32_t* scol, double* sval, uint8_t* skeep, double* sdx, double* sdy, const double* ltheta,
                      hipStream_t s);
hipError_t prune_count(int n, const int64_t* slot_rp, const uint8_t* skeep, int64_t* cnt, int64_t* out_rp,
                       hipStream_t s);
hipError_t prune_compact(int n, int is_complex, const int64_t* slot_rp, const uint8_t* skeep, const int32_t* scol,
                         const double* sval, const int64_t* out_rp, int32_t* ocol, double* oval, hipStream_t s);

}  // namespace leo2d
