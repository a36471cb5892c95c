/*
 * leo2d_oracle.c -- CPU FP64 restatement of LEO2D's KPM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product never
 * does.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no FMA), so every
 * floating-point expression below is evaluated in the reference's order.
 *
 * Reference files cited as path:line relative to the reference tree.
 */
#define _GNU_SOURCE
#include "leo2d_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* constants: src/tools/numbers.h */
#define N_PI 3.14159265358979323846
#define N_PI_6 0.523598775598298873077
#define N_SQRT3_2 0.866025403784438646764
#define GR_ACC 1.429981146728865 /* a / sqrt(3), the unstrained nearest-neighbour bond (A) */
/* tools/numbers.h */
#define ORC_PI 3.14159265358979323846
#define ORC_PI_2 1.57079632679489661923
#define ORC_PI_6 0.523598775598298873077
#define N_SQRT3_6 0.288675134594812882255

/* graphene geometry: src/materials/graphene.h:78-95 */
#define GR_A 2.4768
#define GR_INTER_CUTOFF 8.0
#define GR_INTRA_SEARCH 4
#define GR_INTER_SEARCH 6
/* hstruct.cpp:310 -- pairs farther than this (in-plane) are never stored */
#define INTER_PAIR_CUTOFF 12.5

struct orc_geom {
    orc_geom_desc d;
    int H[ORC_MAX_SHEETS], W[ORC_MAX_SHEETS], count[ORC_MAX_SHEETS], base[ORC_MAX_SHEETS];
    int norb;
    int* grid[ORC_MAX_SHEETS]; /* [H][W][norb], -1 = empty (sheet.cpp:217-255) */
    double lat[2][2];          /* Graphene::lattice */
    double apos[2][3];         /* Graphene::atom_pos */
    double ainv[2][2];         /* Sheet::setInverse, sheet.cpp:887-904 */
    double scinv[ORC_MAX_SHEETS][2][2]; /* Sheet::setSupercell, sheet.cpp:906-924 */
    int n;
    int32_t* i2g;  /* [4n] (i, j, l, s): locality.cpp:1090-1095 */
    double* i2p;   /* [3n] global frame: Hstruct::posAtomIndex, hstruct.cpp:150-218 */
    double* lpos;  /* [3n] sheet-local frame: Sheet::posAtomIndex, sheet.cpp:458-531 */
};

static void graphene_lattice(int mat, double lat[2][2], double apos[2][3])
{
    /* graphene.h:78-85; StrainedGraphene (strained_graphene.h:50-56) rounds sqrt(3)/2
     * and sqrt(3)/6 to 11 digits */
    const double s32 = mat == ORC_MAT_STRAINED_GRAPHENE ? 0.86602540378 : N_SQRT3_2;
    const double s36 = mat == ORC_MAT_STRAINED_GRAPHENE ? 0.28867513459 : N_SQRT3_6;
    lat[0][0] = 1. * GR_A; lat[0][1] = 0.;
    lat[1][0] = .5 * GR_A; lat[1][1] = s32 * GR_A;
    apos[0][0] = 0.; apos[0][1] = 0.; apos[0][2] = 0.;
    apos[1][0] = 0.5 * GR_A; apos[1][1] = s36 * GR_A; apos[1][2] = 0;
}

/* Materials::intra_search_radius / inter_search_radius truncated to int
 * (sheet.cpp:39, 626; hstruct.cpp:305): graphene 4 and 6, StrainedGraphene
 * (5.7 + a/3) and (8 + a/3) */
static int intra_window(int mat) { return mat == ORC_MAT_STRAINED_GRAPHENE ? (int)(5.7 + GR_A / 3.) : GR_INTRA_SEARCH; }
static int inter_window(int mat) { return mat == ORC_MAT_STRAINED_GRAPHENE ? (int)(8. + GR_A / 3.) : GR_INTER_SEARCH; }

/* Sheet::posAtomGrid (sheet.cpp:482-531), solver_space = 0 */
static void pos_atom_grid(const orc_geom* g, int s, int gi, int gj, int l, double p[3])
{
    const orc_sheet_desc* sd = &g->d.sheet[s];
    int i = gi + sd->min_shape[0];
    int j = gj + sd->min_shape[1];
    double x = i * g->lat[0][0] + j * g->lat[1][0] + g->apos[l][0];
    double y = i * g->lat[0][1] + j * g->lat[1][1] + g->apos[l][1];
    double z = g->apos[l][2];
    if (sd->global_shifts_on == 1) {
        x = x + sd->global_shift[0] * g->lat[0][0] + sd->global_shift[1] * g->lat[1][0];
        y = y + sd->global_shift[0] * g->lat[0][1] + sd->global_shift[1] * g->lat[1][1];
    }
    p[0] = x; p[1] = y; p[2] = z;
}

/* Sheet::checkShape (sheet.cpp:537-559) */
static int check_shape(const orc_geom* g, int s, const double p[3])
{
    const orc_sheet_desc* sd = &g->d.sheet[s];
    if (sd->boundary_condition == 1) {
        double delta = 0.0, grid_delta = 0.000001;
        double i = (p[0] + delta) * g->scinv[s][0][0] + (p[1] + delta) * g->scinv[s][0][1];
        double j = (p[0] + delta) * g->scinv[s][1][0] + (p[1] + delta) * g->scinv[s][1][1];
        return (i >= 0.0 - grid_delta && i < 1.0 - grid_delta && j >= 0.0 - grid_delta && j < 1.0 - grid_delta);
    }
    return (p[0] * p[0] + p[1] * p[1] < sd->max_R * sd->max_R);
}

static int sheet_grid(const orc_geom* g, int s, int i, int j, int l)
{
    /* Sheet::gridToIndex, sheet.cpp:572-583 */
    if (i >= 0 && i < g->H[s] && j >= 0 && j < g->W[s] && l >= 0 && l < g->norb)
        return g->grid[s][((size_t)i * g->W[s] + j) * g->norb + l];
    return -1;
}

orc_geom* orc_geom_build(const orc_geom_desc* d)
{
    if (d->num_sheets < 1 || d->num_sheets > ORC_MAX_SHEETS) return NULL;
    orc_geom* g = (orc_geom*)calloc(1, sizeof(orc_geom));
    g->d = *d;
    g->norb = 2;
    for (int s = 1; s < d->num_sheets; ++s)
        if ((d->sheet[s].mat == ORC_MAT_STRAINED_GRAPHENE) != (d->sheet[0].mat == ORC_MAT_STRAINED_GRAPHENE)) {
            free(g);
            return NULL;   /* one lattice for all sheets: StrainedGraphene on every sheet or on none */
        }
    graphene_lattice(d->sheet[0].mat, g->lat, g->apos);
    {
        double a11 = g->lat[0][0], a12 = g->lat[0][1], a21 = g->lat[1][0], a22 = g->lat[1][1];
        double det_a = a11 * a22 - a12 * a21;
        g->ainv[0][0] = a22 / det_a;
        g->ainv[0][1] = -a21 / det_a;
        g->ainv[1][0] = -a12 / det_a;
        g->ainv[1][1] = a11 / det_a;
    }
    int total = 0;
    for (int s = 0; s < d->num_sheets; ++s) {
        const orc_sheet_desc* sd = &d->sheet[s];
        if (sd->boundary_condition == 1) {
            const double(*sc)[2] = sd->supercell;
            double det = sc[0][0] * sc[1][1] - sc[0][1] * sc[1][0];
            g->scinv[s][0][0] = sc[1][1] / det;
            g->scinv[s][0][1] = -sc[1][0] / det;
            g->scinv[s][1][0] = -sc[0][1] / det;
            g->scinv[s][1][1] = sc[0][0] / det;
        }
        /* Sheet::setIndex, sheet.cpp:217-285: i -> j -> l order */
        int H = sd->max_shape[0] - sd->min_shape[0];
        int W = sd->max_shape[1] - sd->min_shape[1];
        if (H < 0) H = 0;
        if (W < 0) W = 0;
        g->H[s] = H; g->W[s] = W;
        g->grid[s] = (int*)malloc(sizeof(int) * ((size_t)H * W * g->norb + 1));
        int k = 0;
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j)
                for (int l = 0; l < g->norb; ++l) {
                    double p[3];
                    pos_atom_grid(g, s, i, j, l, p);
                    g->grid[s][((size_t)i * W + j) * g->norb + l] = check_shape(g, s, p) ? k++ : -1;
                }
        g->count[s] = k;
        g->base[s] = total;
        total += k;
    }
    g->n = total;
    g->i2g = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(total + 1));
    g->i2p = (double*)malloc(sizeof(double) * 3 * (size_t)(total + 1));
    g->lpos = (double*)malloc(sizeof(double) * 3 * (size_t)(total + 1));
    for (int s = 0; s < d->num_sheets; ++s) {
        double theta = d->sheet[s].angle;
        for (int i = 0; i < g->H[s]; ++i)
            for (int j = 0; j < g->W[s]; ++j)
                for (int l = 0; l < g->norb; ++l) {
                    int k = g->grid[s][((size_t)i * g->W[s] + j) * g->norb + l];
                    if (k < 0) continue;
                    size_t K = (size_t)(k + g->base[s]);
                    g->i2g[4 * K + 0] = i; g->i2g[4 * K + 1] = j;
                    g->i2g[4 * K + 2] = l; g->i2g[4 * K + 3] = s;
                    double p[3];
                    pos_atom_grid(g, s, i, j, l, p);
                    g->lpos[3 * K + 0] = p[0]; g->lpos[3 * K + 1] = p[1]; g->lpos[3 * K + 2] = p[2];
                    /* Hstruct::posAtomIndex (hstruct.cpp:206-217); Hstruct shifts are zero */
                    double zero = 0.0;
                    g->i2p[3 * K + 0] = p[0] * cos(theta) - p[1] * sin(theta) + zero * g->lat[0][0] + zero * g->lat[1][0];
                    g->i2p[3 * K + 1] = p[0] * sin(theta) + p[1] * cos(theta) + zero * g->lat[0][1] + zero * g->lat[1][1];
                    g->i2p[3 * K + 2] = p[2] + d->sheet[s].height;
                }
    }
    return g;
}

void orc_geom_free(orc_geom* g)
{
    if (!g) return;
    for (int s = 0; s < g->d.num_sheets; ++s) free(g->grid[s]);
    free(g->i2g); free(g->i2p); free(g->lpos);
    free(g);
}

int orc_geom_n(const orc_geom* g) { return g->n; }
int orc_geom_sheet_count(const orc_geom* g, int s) { return g->count[s]; }

void orc_geom_arrays(const orc_geom* g, int32_t* index_to_grid, double* index_to_pos)
{
    if (index_to_grid) memcpy(index_to_grid, g->i2g, sizeof(int32_t) * 4 * (size_t)g->n);
    if (index_to_pos) memcpy(index_to_pos, g->i2p, sizeof(double) * 3 * (size_t)g->n);
}

int orc_grid_to_index(const orc_geom* g, int i, int j, int l, int s)
{
    /* Hstruct::gridToIndex (hstruct.cpp:821-848) */
    if (s < 0 || s >= g->d.num_sheets) return -1;
    int k = sheet_grid(g, s, i, j, l);
    if (k == -1) return -1;
    return k + g->base[s];
}

int orc_target_center(const orc_geom* g, int s)
{
    /* Hstruct::getTargetList (hstruct.cpp:1063-1070): centre unit cell, orbital 0 */
    const orc_sheet_desc* sd = &g->d.sheet[s];
    int cx = (sd->max_shape[0] - sd->min_shape[0]) / 2;
    int cy = (sd->max_shape[1] - sd->min_shape[1]) / 2;
    return orc_grid_to_index(g, cx, cy, 0, s);
}

/* ------------------------------------------------------------------ */
/* couplings                                                            */

static double graphene_intra(int orb_row, int orb_col, int v0, int v1)
{
    /* Coupling::Intralayer::graphene, graphene.cpp:13-89 (orbital 0 = A_pz, 1 = B_pz) */
    int h0 = 2 * v0 + v1;
    int h1 = -1 * v0 + v1;
    if (orb_col == 0 && orb_row == 1) h0 -= 1;
    else if (orb_col == 1 && orb_row == 0) h0 += 1;
    int r = h0 * (h0 + h1) + h1 * h1;
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
    default: return 0.;
    }
}

static double simple_intra(int orb_row, int orb_col, int v0, int v1)
{
    /* nearest-neighbour-only model of the workload configs: same honeycomb
     * shell index as graphene.cpp:40-61, t = -2.7 eV on shell 1, 0 elsewhere */
    int h0 = 2 * v0 + v1;
    int h1 = -1 * v0 + v1;
    if (orb_col == 0 && orb_row == 1) h0 -= 1;
    else if (orb_col == 1 && orb_row == 0) h0 += 1;
    int r = h0 * (h0 + h1) + h1 * h1;
    return r == 1 ? -2.7 : 0.0;
}

/* honeycomb shell index of a hop (strained_graphene.cpp:16-60, the same homogeneous
 * coordinates as graphene.cpp:40-61) */
static int hop_shell(int orb_row, int orb_col, int v0, int v1)
{
    int h0 = 2 * v0 + v1;
    int h1 = -1 * v0 + v1;
    if (orb_col == 0 && orb_row == 1) h0 -= 1;
    else if (orb_col == 1 && orb_row == 0) h0 += 1;
    return h0 * (h0 + h1) + h1 * h1;
}

static const double SG_T[4] = {0.7833, -2.8219, 0.2543, -0.1803};   /* strained_graphene.cpp:81 */
static const double SG_A[4] = {-4.8782, 4.0066, -0.4633, 0.6236};
static const double SG_B[4] = {0.0000, -3.0868, 0.8017, 0.4793};

static double strained_table(int orb_row, int orb_col, int v0, int v1)
{
    /* strained_graphene.cpp:16-91: shells 0, 1, 3, 4 */
    switch (hop_shell(orb_row, orb_col, v0, v1)) {
    case 0: return SG_T[0];
    case 1: return SG_T[1];
    case 3: return SG_T[2];
    case 4: return SG_T[3];
    default: return 0.;
    }
}

double orc_strained_intra(int orb_row, int orb_col, int v0, int v1, double hx, double hy, const double u[2][2],
                          double theta)
{
    /* strained_graphene.cpp:300-396.  The second component of the vector-like strain
     * is written twice there (u_xx - u_yy, then u_xy + u_yx into the same slot), so
     * the pair is (u_xy + u_yx, 0). */
    const double onedim = u[0][0] + u[1][1];
    const double two0 = u[0][1] + u[1][0], two1 = 0.0;
    const double w0 = cos(theta) * hx + sin(theta) * hy;
    const double w1 = -sin(theta) * hx + cos(theta) * hy;
    switch (hop_shell(orb_row, orb_col, v0, v1)) {
    case 0: return SG_T[0] + SG_A[0] * onedim;
    case 1: return SG_T[1] + SG_A[1] * onedim + SG_B[1] * (w0 * two0 + w1 * two1);
    case 3: return SG_T[2] + SG_A[2] * onedim + SG_B[2] * (-w1 * two0 + w0 * two1);
    case 4: return SG_T[3] + SG_A[3] * onedim + SG_B[3] * (w0 * two0 + w1 * two1);
    default: return 0.;
    }
}

double orc_intralayer(int mat, int orb_row, int orb_col, int v0, int v1)
{
    if (mat == ORC_MAT_GRAPHENE) return graphene_intra(orb_row, orb_col, v0, v1);
    if (mat == ORC_MAT_STRAINED_GRAPHENE) return strained_table(orb_row, orb_col, v0, v1);
    return simple_intra(orb_row, orb_col, v0, v1);
}

static double c_to_c(int orb_row, int orb_col, double v0, double v1, double v2, double theta_row, double theta_col)
{
    /* Coupling::Interlayer::C_to_C, graphene.cpp:91-222.  The |z| test at
     * graphene.cpp:117 calls the C library's int abs() on a double (the
     * argument is truncated toward zero); reproduced as such. */
    if (abs((int)v2) > 5) return 0;
    double r = sqrt(v0 * v0 + v1 * v1);
    if (r < GR_INTER_CUTOFF) {
        double ac = atan2(v1, v0);
        double theta21 = ac - theta_row;
        if (orb_row == 1) theta21 += N_PI_6;
        else theta21 -= N_PI_6;
        double theta12 = ac - theta_col + N_PI;
        if (orb_col == 1) theta12 += N_PI_6;
        else theta12 -= N_PI_6;
        double rs = r / GR_A;
        double z_eq = 3.35;
        double z_eps = (fabs(v2) - z_eq) / z_eq;
        double lambda_0 = 0.310 - 1.882 * z_eps + 7.741 * z_eps * z_eps;
        double xi_0 = 1.750 - 1.618 * z_eps + 1.848 * z_eps * z_eps;
        double kappa_0 = 1.990 + 1.007 * z_eps + 2.427 * z_eps * z_eps;
        double lambda_3 = -0.068 + 0.399 * z_eps - 1.739 * z_eps * z_eps;
        double xi_3 = 3.286 - 0.914 * z_eps + 12.011 * z_eps * z_eps;
        double x_3 = 0.500 + 0.322 * z_eps + 0.908 * z_eps * z_eps;
        double lambda_6 = -0.008 + 0.046 * z_eps - 0.183 * z_eps * z_eps;
        double xi_6 = 2.272 - 0.721 * z_eps - 4.414 * z_eps * z_eps;
        double x_6 = 1.217 + 0.027 * z_eps - 0.658 * z_eps * z_eps;
        double kappa_6 = 1.562 - 0.371 * z_eps - 0.134 * z_eps * z_eps;
        double V0 = lambda_0 * exp(-xi_0 * (rs) * (rs)) * cos(kappa_0 * rs);
        double V3 = lambda_3 * rs * rs * exp(-xi_3 * (rs - x_3) * (rs - x_3));
        double V6 = lambda_6 * exp(-xi_6 * (rs - x_6) * (rs - x_6)) * sin(kappa_6 * rs);
        double t = V0 + V3 * (cos(3 * theta12) + cos(3 * theta21)) + V6 * (cos(6 * theta12) + cos(6 * theta21));
        double d = GR_INTER_CUTOFF - r;
        if (d < 1. && d > 0.) t *= exp(1. - 1. / (d * d));
        return t;
    }
    return 0.;
}

static double simple_inter(double v0, double v1, double v2)
{
    /* t_perp(r) = 0.48 eV * exp(-(r - 3.35 A)/0.45 A) for 3D distance r < 6 A */
    double r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (r < 6.0) return 0.48 * exp(-(r - 3.35) / 0.45);
    return 0.0;
}

double orc_interlayer(int mat, int orb_row, int orb_col, double dx, double dy, double dz,
                      double theta_row, double theta_col)
{
    /* Materials::interlayer_term (materials.cpp:362-372): C_to_C for Graphene and StrainedGraphene */
    return (mat == ORC_MAT_GRAPHENE || mat == ORC_MAT_STRAINED_GRAPHENE)
               ? c_to_c(orb_row, orb_col, dx, dy, dz, theta_row, theta_col)
               : simple_inter(dx, dy, dz);
}

/* ------------------------------------------------------------------ */
/* pair lists                                                           */

typedef struct { int32_t i, j; int32_t sc0, sc1; double t; } pair_t;

struct orc_pairs {
    int64_t np[2];
    pair_t* p[2];
};

typedef struct { pair_t* a; int64_t n, cap; } pvec;
static void pv_push(pvec* v, pair_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->a = (pair_t*)realloc(v->a, sizeof(pair_t) * (size_t)v->cap);
    }
    v->a[v->n++] = x;
}

/* Sheet::orderPairs / Hstruct::orderPairs (sheet.cpp:810-849, hstruct.cpp:634-666)
 * are bubble sorts on the column with a strict '>' swap test, i.e. a stable
 * sort by column.  Restated as a stable insertion sort (same permutation). */
static void stable_sort_by_col(pair_t* a, int64_t n)
{
    for (int64_t i = 1; i < n; ++i) {
        pair_t x = a[i];
        int64_t j = i - 1;
        while (j >= 0 && a[j].j > x.j) { a[j + 1] = a[j]; --j; }
        a[j + 1] = x;
    }
}

/* Sheet::findNearest (sheet.cpp:856-882) */
static int sheet_nearest(const orc_geom* g, int s, const double p[3], int dim)
{
    const orc_sheet_desc* sd = &g->d.sheet[s];
    double v = g->ainv[dim][0] * p[0] + g->ainv[dim][1] * p[1];
    int f = (int)floor(v);
    if (f < sd->min_shape[dim]) f = sd->min_shape[dim];
    if (f > sd->max_shape[dim]) f = sd->max_shape[dim];
    return f - sd->min_shape[dim];
}

/* Hstruct::findNearest (hstruct.cpp:224-288), solver_space = 0, zero Hstruct shifts */
static int hstruct_nearest(const orc_geom* g, const double p[3], int s, int dim)
{
    const orc_sheet_desc* sd = &g->d.sheet[s];
    double theta = -sd->angle;
    double zero = 0.0;
    double x = p[0], y = p[1];
    double x_new = x * cos(theta) - y * sin(theta) + zero * g->lat[0][0] + zero * g->lat[1][0];
    double y_new;
    if (dim == 0) y_new = x * sin(theta) + y * cos(theta) + zero * g->lat[0][1] + zero * g->lat[1][1];
    else y_new = y * cos(theta) + x * sin(theta) + zero * g->lat[0][1] + zero * g->lat[1][1];
    double v = g->ainv[dim][0] * x_new + g->ainv[dim][1] * y_new;
    int f = (int)floor(v);
    if (f < sd->min_shape[dim]) f = sd->min_shape[dim];
    if (f > sd->max_shape[dim]) f = sd->max_shape[dim];
    return f - sd->min_shape[dim];
}

static void intra_pairs_sheet(const orc_geom* g, int s, pvec* out)
{
    /* Sheet::getIntraPairs (sheet.cpp:624-765), solver_space = 0, mat_from_file = 0 */
    const orc_sheet_desc* sd = &g->d.sheet[s];
    int S = sd->boundary_condition == 1 ? g->d.sc_search_size : 0;
    int start = g->base[s];
    pvec row = {0};
    for (int kh = 0; kh < g->count[s]; ++kh) {
        row.n = 0;
        size_t K = (size_t)(kh + start);
        int ih = g->i2g[4 * K + 0], jh = g->i2g[4 * K + 1], l0 = g->i2g[4 * K + 2];
        const double* kp = &g->lpos[3 * K];
        for (int dx = -S; dx < S + 1; ++dx)
            for (int dy = -S; dy < S + 1; ++dy) {
                double scd0 = 0.0, scd1 = 0.0;
                if (sd->boundary_condition == 1) {
                    scd0 = dx * sd->supercell[0][0] + dy * sd->supercell[1][0];
                    scd1 = dx * sd->supercell[0][1] + dy * sd->supercell[1][1];
                }
                double ph[3] = {kp[0] + scd0, kp[1] + scd1, kp[2]};
                int i0 = sheet_nearest(g, s, ph, 0);
                int j0 = sheet_nearest(g, s, ph, 1);
                const int W = intra_window(sd->mat);
                for (int i = i0 - W; i < i0 + W; ++i)
                    for (int j = j0 - W; j < j0 + W; ++j)
                        for (int l = 0; l < g->norb; ++l) {
                            int k2 = sheet_grid(g, s, i, j, l);
                            if (k2 == -1) continue;
                            int gd0 = i - ih, gd1 = j - jh;
                            if (sd->boundary_condition == 1) {
                                gd0 = gd0 - dx * sd->supercell_stride[0][0] - dy * sd->supercell_stride[1][0];
                                gd1 = gd1 - dx * sd->supercell_stride[0][1] - dy * sd->supercell_stride[1][1];
                            }
                            double t = orc_intralayer(sd->mat, l0, l, gd0, gd1);
                            if (t != 0 || (dx == 0 && dy == 0)) {
                                pair_t pr = {kh + start, k2 + start, dx, dy, t};
                                pv_push(&row, pr);
                            }
                        }
            }
        if (sd->boundary_condition == 1) stable_sort_by_col(row.a, row.n);
        for (int64_t q = 0; q < row.n; ++q) pv_push(out, row.a[q]);
    }
    free(row.a);
}

static void inter_pairs_all(const orc_geom* g, pvec* out)
{
    /* Hstruct::getInterPairs (hstruct.cpp:298-632), solver_space = 0 */
    int bc = g->d.boundary_condition;
    int S = bc == 1 ? g->d.sc_search_size : 0;
    const double cut2 = INTER_PAIR_CUTOFF * INTER_PAIR_CUTOFF;
    pvec row = {0};
    for (int kh = 0; kh < g->n; ++kh) {
        row.n = 0;
        int sh = g->i2g[4 * (size_t)kh + 3];
        const double* kp = &g->i2p[3 * (size_t)kh];
        for (int dx = -S; dx < S + 1; ++dx)
            for (int dy = -S; dy < S + 1; ++dy) {
                double scd0 = 0.0, scd1 = 0.0;
                if (bc == 1) {
                    scd0 = dx * g->d.supercell[0][0] + dy * g->d.supercell[1][0];
                    scd1 = dx * g->d.supercell[0][1] + dy * g->d.supercell[1][1];
                }
                double ph[3] = {kp[0] + scd0, kp[1] + scd1, kp[2]};
                for (int side = 0; side < 2; ++side) {
                    int s0;
                    if (side == 0) { if (!(sh > 0)) continue; s0 = sh - 1; }
                    else { if (!(sh < g->d.num_sheets - 1)) continue; s0 = sh + 1; }
                    int i0 = hstruct_nearest(g, ph, s0, 0);
                    int j0 = hstruct_nearest(g, ph, s0, 1);
                    int base = g->base[s0];
                    const int W = inter_window(g->d.sheet[0].mat);
                    int ilo = i0 - W > 0 ? i0 - W : 0;
                    int ihi = g->H[s0] < i0 + W ? g->H[s0] : i0 + W;
                    int jlo = j0 - W > 0 ? j0 - W : 0;
                    int jhi = g->W[s0] < j0 + W ? g->W[s0] : j0 + W;
                    for (int i = ilo; i < ihi; ++i)
                        for (int j = jlo; j < jhi; ++j)
                            for (int l = 0; l < g->norb; ++l) {
                                int k2 = sheet_grid(g, s0, i, j, l);
                                if (k2 == -1) continue;
                                double x2 = g->i2p[3 * (size_t)(k2 + base) + 0];
                                double y2 = g->i2p[3 * (size_t)(k2 + base) + 1];
                                if ((x2 - ph[0]) * (x2 - ph[0]) + (y2 - ph[1]) * (y2 - ph[1]) < cut2) {
                                    pair_t pr = {kh, k2 + base, dx, dy, 0.0};
                                    pv_push(&row, pr);
                                }
                            }
                }
            }
        if (bc == 1) stable_sort_by_col(row.a, row.n);
        for (int64_t q = 0; q < row.n; ++q) pv_push(out, row.a[q]);
    }
    free(row.a);
}

orc_pairs* orc_pairs_build(const orc_geom* g)
{
    orc_pairs* p = (orc_pairs*)calloc(1, sizeof(orc_pairs));
    pvec intra = {0}, inter = {0};
    /* Hstruct::getIntraPairs (hstruct.cpp:886-893): sheet by sheet */
    for (int s = 0; s < g->d.num_sheets; ++s) intra_pairs_sheet(g, s, &intra);
    inter_pairs_all(g, &inter);
    p->np[0] = intra.n; p->p[0] = intra.a;
    p->np[1] = inter.n; p->p[1] = inter.a;
    return p;
}

void orc_pairs_free(orc_pairs* p)
{
    if (!p) return;
    free(p->p[0]); free(p->p[1]); free(p);
}

int64_t orc_pairs_count(const orc_pairs* p, int which) { return p->np[which]; }

void orc_pairs_copy(const orc_pairs* p, int which, int32_t* ij, int32_t* sc, double* t)
{
    for (int64_t q = 0; q < p->np[which]; ++q) {
        const pair_t* x = &p->p[which][q];
        if (ij) { ij[2 * q] = x->i; ij[2 * q + 1] = x->j; }
        if (sc) { sc[2 * q] = x->sc0; sc[2 * q + 1] = x->sc1; }
        if (t) t[q] = x->t;
    }
}

/* ------------------------------------------------------------------ */
/* assembly                                                             */

void orc_relax_disp(const orc_job* job, double x, double y, int s, double* d)
{
    /* sum over Fourier directions k = i b1 + j b2 (j < max(i, 1)), each with its
     * three 60-degree rotations (one for the constant term), coefficients in
     * the supercell basis; then turned by the angle of b1 (strain.cpp:390-460)
     * and given the sheet's sign: +1, or -1 on sheet 2 (strain.cpp:290-311).
     * STRAIN_TYPE = DIRECT (6, StrainCalc::supercellDisp, strain.cpp:756-876) uses
     * the same sum over fixed max_k = 2 coefficients of a relaxed graphene
     * sandwich, scaled per sheet before the sum (in plane x -2 on sheet 2, out of
     * plane x 0 on sheet 2 and x -1 on sheet 3) and no sign. */
    const double(*b)[2] = job->relax_b;
    int max_k = job->relax_max_k;
    const double* cx = NULL;
    const double* cy = NULL;
    const double* cz = NULL;
    double c6[3][4];
    if (job->strain_type == 6) {
        static const double sx[4] = {0.0, 0.0, 0.0, 0.0};
        static const double sy[4] = {0.0, -0.019951155162081, -0.000806807799645, 0.0};
        static const double sz[4] = {-0.061901691526695, -0.0220634404118861, -0.004100201789620, -0.002676762110702};
        const double fin = s == 1 ? -2.0 : 1.0;
        const double fout = s == 2 ? -1.0 : (s == 1 ? 0.0 : 1.0);
        for (int q = 0; q < 4; ++q) {
            c6[0][q] = fin * sx[q];
            c6[1][q] = fin * sy[q];
            c6[2][q] = fout * sz[q];
        }
        max_k = 2;
        cx = c6[0]; cy = c6[1]; cz = c6[2];
    } else {
        const int nc = max_k * (max_k + 1) / 2 + 1;
        cx = job->relax;
        cy = job->relax + nc;
        cz = job->relax + 2 * nc;
    }
    double ux = 0.0, uy = 0.0, uz = 0.0;
    int idx = 0;
    for (int i = 0; i < max_k + 1; ++i) {
        for (int j = 0; j < (i > 1 ? i : 1); ++j) {
            const double ff = i == 0 ? 1.0 : 2.0;
            const int rot_max = i == 0 ? 1 : 3;
            const double kb0 = i * b[0][0] + j * b[1][0], kb1 = i * b[0][1] + j * b[1][1];
            for (int r = 0; r < rot_max; ++r) {
                const double th = r * ORC_PI / 3.0;
                const double k0 = cos(th) * kb0 - sin(th) * kb1, k1 = sin(th) * kb0 + cos(th) * kb1;
                const double c0 = cos(th) * cx[idx] - sin(th) * cy[idx], c1 = sin(th) * cx[idx] + cos(th) * cy[idx];
                ux = ux + ff * c0 * sin(k0 * x + k1 * y);
                uy = uy + ff * c1 * sin(k0 * x + k1 * y);
                uz = uz + ff * cz[idx] * cos(k0 * x + k1 * y);
            }
            ++idx;
        }
    }
    const double rot = atan2(b[0][1], b[0][0]);
    const double rx = cos(rot) * ux - sin(rot) * uy, ry = sin(rot) * ux + cos(rot) * uy;
    if (job->strain_type == 6) {
        d[0] = rx;
        d[1] = ry;
        d[2] = uz;
        return;
    }
    const double sign = s == 1 ? -1.0 : 1.0;
    d[0] = sign * rx;
    d[1] = sign * ry;
    d[2] = sign * uz;
}

void orc_local_theta(const orc_geom* g, const double* i2pos, double* theta)
{
    for (int k1 = 0; k1 < g->n; ++k1) {
        const int* a = g->i2g + 4 * (size_t)k1;
        int k2 = -1;
        for (int dk = -3; dk < 4 && k2 < 0; dk += 2) {
            const int t = k1 + dk;
            if (t < 0 || t >= g->n) continue;
            const int* c = g->i2g + 4 * (size_t)t;
            /* isNearestNeighbor (locality.cpp:733-767): same sheet, the other
             * orbital, same cell or the cell at -1 (from A) / +1 (from B) in i or j */
            if (c[3] != a[3] || c[2] == a[2]) continue;
            const int di = c[0] - a[0], dj = c[1] - a[1];
            const int sgn = a[2] == 0 ? -1 : 1;
            if ((di == 0 && dj == 0) || (di == sgn && dj == 0) || (di == 0 && dj == sgn)) k2 = t;
        }
        if (k2 >= 0) {
            const double dx = i2pos[3 * (size_t)k2] - i2pos[3 * (size_t)k1];
            const double dy = i2pos[3 * (size_t)k2 + 1] - i2pos[3 * (size_t)k1 + 1];
            theta[k1] = atan2(dy, dx) - (a[2] == 0 ? ORC_PI_6 : ORC_PI_2);
        } else {
            theta[k1] = g->d.sheet[a[3]].angle;
        }
    }
}

void orc_job_positions(const orc_geom* g, const orc_job* job, double* i2pos)
{
    /* Locality::setConfigPositions (locality.cpp:3618-3643, 3849-3857) */
    for (int i = 0; i < g->n; ++i) {
        int s = g->i2g[4 * (size_t)i + 3];
        double theta = g->d.sheet[s].angle;
        double s1_a[2][2];
        for (int dd = 0; dd < 2; ++dd) {
            s1_a[dd][0] = cos(theta) * g->lat[dd][0] - sin(theta) * g->lat[dd][1];
            s1_a[dd][1] = sin(theta) * g->lat[dd][0] + cos(theta) * g->lat[dd][1];
        }
        const double* sh = job->shift[s];
        i2pos[i * 3 + 0] = g->i2p[i * 3 + 0] + sh[0] * s1_a[0][0] + sh[1] * s1_a[1][0];
        i2pos[i * 3 + 1] = g->i2p[i * 3 + 1] + sh[0] * s1_a[0][1] + sh[1] * s1_a[1][1];
        i2pos[i * 3 + 2] = g->i2p[i * 3 + 2];
        if (job->strain_type == 1 || job->strain_type == 6) {
            /* locality.cpp:3657-3711 */
            double d[3];
            orc_relax_disp(job, i2pos[i * 3 + 0], i2pos[i * 3 + 1], s, d);
            i2pos[i * 3 + 0] = i2pos[i * 3 + 0] + d[0];
            i2pos[i * 3 + 1] = i2pos[i * 3 + 1] + d[1];
            i2pos[i * 3 + 2] = i2pos[i * 3 + 2] + d[2];
        }
        if (job->sheet_strain_on == 1) {
            const double(*e)[2] = job->sheet_strain[s];
            double tx = i2pos[i * 3 + 0], ty = i2pos[i * 3 + 1];
            i2pos[i * 3 + 0] = (1.0 + e[0][0]) * tx + e[0][1] * ty;
            i2pos[i * 3 + 1] = e[1][0] * tx + (1.0 + e[1][1]) * ty;
        }
        if (job->uniform_strain == 1) {
            double tx = i2pos[i * 3 + 0], ty = i2pos[i * 3 + 1];
            i2pos[i * 3 + 0] = (1.0 + job->u[0][0]) * tx + job->u[0][1] * ty;
            i2pos[i * 3 + 1] = job->u[1][0] * tx + (1.0 + job->u[1][1]) * ty;
        }
    }
}

/* support of a coupling: used only by the pruned pattern mode (an extension:
 * the reference stores every candidate slot).  Geometric so the predicate is
 * decided by IEEE-exact operations on both the CPU and the GPU. */
static int inter_in_support(int mat, double dx, double dy, double dz)
{
    if (mat != ORC_MAT_GRAPHENE && mat != ORC_MAT_STRAINED_GRAPHENE) return sqrt(dx * dx + dy * dy + dz * dz) < 6.0;
    if (abs((int)dz) > 5) return 0;
    return sqrt(dx * dx + dy * dy) < GR_INTER_CUTOFF;
}

int orc_assemble(const orc_geom* g, const orc_pairs* p, const orc_job* job, orc_csr* out)
{
    /* Locality::generateRealH (locality.cpp:3900-4321) and generateCpxH
     * (locality.cpp:4485-5158), strain_type = 0, no vacancies. */
    int n = g->n;
    int cpx = job->complex_matrix;
    double* i2pos = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    orc_job_positions(g, job, i2pos);
    /* with a relaxation field the interlayer term takes each orbital's local twist
     * (locality.cpp:4261-4269, 5045-5052) */
    double* ltheta = NULL;
    if (job->strain_type != 0) {
        ltheta = (double*)malloc(sizeof(double) * (size_t)(n + 1));
        orc_local_theta(g, i2pos, ltheta);
    }
    double sc[2][2];
    memcpy(sc, g->d.supercell, sizeof(sc));
    if (g->d.boundary_condition == 1 && job->uniform_strain == 1) {
        /* locality.cpp:2947-2960 */
        for (int i = 0; i < 2; ++i) {
            sc[i][0] = (1.0 + job->u[0][0]) * g->d.supercell[i][0] + job->u[0][1] * g->d.supercell[i][1];
            sc[i][1] = job->u[1][0] * g->d.supercell[i][0] + (1.0 + job->u[1][1]) * g->d.supercell[i][1];
        }
    }
    int64_t maxnnz = p->np[0] + p->np[1];
    int64_t* rp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxnnz + 1));
    double* v = (double*)calloc((size_t)(cpx ? 2 : 1) * (size_t)(maxnnz + 1), sizeof(double));
    /* dxH slots (locality.cpp:4182, 4299): (x1 - x2) t accumulated with H */
    /* velocity slots (locality.cpp:4182-4183, 4299-4300 real; 4939-4940, 5113-5114 complex):
     * dxH, dyH accumulate (x1 - x2) t and (y1 - y2) t on H's slots */
    const int w = cpx ? 2 : 1;
    double* vdx = (double*)calloc((size_t)w * (size_t)(maxnnz + 1), sizeof(double));
    double* vdy = (double*)calloc((size_t)w * (size_t)(maxnnz + 1), sizeof(double));
    char* keep = (char*)calloc((size_t)(maxnnz + 1), 1);
    int64_t ic = 0, xc = 0, input_counter = 0;
    const double resc = job->energy_rescale, eshift = job->energy_shift;
    int bc = g->d.boundary_condition;
    for (int k_i = 0; k_i < n; ++k_i) {
        rp[k_i] = input_counter;
        int s_i = g->i2g[4 * (size_t)k_i + 3];
        int mat = g->d.sheet[s_i].mat;
        for (int blk = 0; blk < 2; ++blk) {
            const pair_t* P = p->p[blk];
            int64_t NP = p->np[blk];
            int64_t* ctr = blk == 0 ? &ic : &xc;
            while (*ctr < NP && P[*ctr].i == k_i) {
                const pair_t* pr = &P[*ctr];
                int new_k = pr->j;
                int sc_i = 0, sc_j = 0;
                double psx = 0.0, psy = 0.0;
                if (bc == 1) {
                    sc_i = pr->sc0; sc_j = pr->sc1;
                    psx = sc_i * sc[0][0] + sc_j * sc[1][0];
                    psy = sc_i * sc[0][1] + sc_j * sc[1][1];
                }
                double x1 = i2pos[k_i * 3 + 0] + psx;
                double y1 = i2pos[k_i * 3 + 1] + psy;
                double z1 = i2pos[k_i * 3 + 2];
                double x2 = i2pos[new_k * 3 + 0];
                double y2 = i2pos[new_k * 3 + 1];
                double z2 = i2pos[new_k * 3 + 2];
                double t;
                int support;
                if (blk == 0) {
                    double raw_t = pr->t;
                    if (mat == ORC_MAT_STRAINED_GRAPHENE &&
                        (job->strain_type == 1 || job->strain_type == 6 || job->uniform_strain == 1)) {
                        /* strained intralayer term (locality.cpp:4034-4160, 4744-4870): the hop's
                         * direction on the job's positions, the uniform strain (zero for the
                         * relaxation field) and the sheet angle */
                        int lh = g->i2g[4 * (size_t)new_k + 2];
                        int gd0 = g->i2g[4 * (size_t)new_k + 0] - g->i2g[4 * (size_t)k_i + 0];
                        int gd1 = g->i2g[4 * (size_t)new_k + 1] - g->i2g[4 * (size_t)k_i + 1];
                        if (bc == 1) {
                            const orc_sheet_desc* sd = &g->d.sheet[s_i];
                            gd0 = gd0 - sc_i * sd->supercell_stride[0][0] - sc_j * sd->supercell_stride[1][0];
                            gd1 = gd1 - sc_i * sd->supercell_stride[0][1] - sc_j * sd->supercell_stride[1][1];
                        }
                        double hx = x2 - x1, hy = y2 - y1;
                        double hn = sqrt(hx * hx + hy * hy);
                        if (hn != 0) { hx = hx / hn; hy = hy / hn; }
                        const double zero_u[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
                        raw_t = orc_strained_intra(g->i2g[4 * (size_t)k_i + 2], lh, gd0, gd1, hx, hy,
                                                   job->uniform_strain == 1 ? job->u : zero_u,
                                                   g->d.sheet[s_i].angle);
                    }
                    if (mat == ORC_MAT_GRAPHENE_SIMPLE_STRAIN && raw_t != 0) {
                        /* strained bond: t = -2.7 exp(-3.37 (d / a_cc - 1)) */
                        double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
                        raw_t = raw_t * exp(-3.37 * (sqrt(dx * dx + dy * dy + dz * dz) / GR_ACC - 1.0));
                    }
                    int diag = cpx ? (new_k == k_i && sc_i == 0 && sc_j == 0) : (new_k == k_i);
                    if (diag) {
                        if (job->elec_on == 1) t = (raw_t + eshift + z1 * job->E) / resc;
                        else t = (raw_t + eshift) / resc;
                    } else {
                        t = raw_t / resc;
                    }
                    support = raw_t != 0 || (diag && (eshift != 0 || job->elec_on == 1));
                } else {
                    int o1 = g->i2g[4 * (size_t)k_i + 2], o2 = g->i2g[4 * (size_t)new_k + 2];
                    int s2 = g->i2g[4 * (size_t)new_k + 3];
                    double th1 = ltheta ? ltheta[k_i] : g->d.sheet[s_i].angle;
                    double th2 = ltheta ? ltheta[new_k] : g->d.sheet[s2].angle;
                    double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
                    t = orc_interlayer(mat, o1, o2, dx, dy, dz, th1, th2) / resc;
                    support = inter_in_support(mat, dx, dy, dz);
                }
                if (!cpx) {
                    if (t != 0) {
                        v[input_counter] += t;
                        vdx[input_counter] += (x1 - x2) * t;
                        vdy[input_counter] += (y1 - y2) * t;
                    }
                } else {
                    /* peierlsPhase (locality.cpp:7157-7161) + k.(r2 - r1) */
                    double phase = 3.14e-5 * job->B * (x2 - x1) * (0.5) * (y1 + y2);
                    phase = phase + job->k[0] * (x2 - x1) + job->k[1] * (y2 - y1);
                    double tr = t * cos(phase), ti = t * sin(phase); /* std::polar */
                    if (tr != 0 || ti != 0) {
                        v[2 * input_counter] += tr;
                        v[2 * input_counter + 1] += ti;
                        vdx[2 * input_counter] += (x1 - x2) * tr;
                        vdx[2 * input_counter + 1] += (x1 - x2) * ti;
                        vdy[2 * input_counter] += (y1 - y2) * tr;
                        vdy[2 * input_counter + 1] += (y1 - y2) * ti;
                    }
                }
                if (support) keep[input_counter] = 1;
                int last = (*ctr + 1 == NP) || P[*ctr + 1].i != k_i || P[*ctr + 1].j != new_k;
                if (last) {
                    ci[input_counter] = new_k;
                    ++input_counter;
                }
                ++*ctr;
            }
        }
    }
    rp[n] = input_counter;
    free(i2pos);
    free(ltheta);
    if (job->pruned) {
        int64_t wr = 0;
        for (int r = 0; r < n; ++r) {
            int64_t b = rp[r], e = rp[r + 1];
            rp[r] = wr;
            for (int64_t q = b; q < e; ++q) {
                if (!keep[q]) continue;
                ci[wr] = ci[q];
                for (int c = 0; c < w; ++c) {
                    v[w * wr + c] = v[w * q + c];
                    vdx[w * wr + c] = vdx[w * q + c];
                    vdy[w * wr + c] = vdy[w * q + c];
                }
                ++wr;
            }
        }
        rp[n] = wr;
    }
    free(keep);
    out->n = n;
    out->nnz = rp[n];
    out->row_ptr = rp;
    out->col = ci;
    out->val = v;
    out->is_complex = cpx;
    out->val_dx = vdx;
    out->val_dy = vdy;
    return 0;
}

void orc_csr_free(orc_csr* h)
{
    free(h->row_ptr); free(h->col); free(h->val); free(h->val_dx); free(h->val_dy);
    h->row_ptr = NULL; h->col = NULL; h->val = NULL; h->val_dx = NULL; h->val_dy = NULL;
}

/* ------------------------------------------------------------------ */
/* KPM                                                                  */

void orc_spmv(int n, const int64_t* rp, const int32_t* ci, const double* val, int is_complex,
              const double* x, double* y, double alpha, double beta)
{
    /* SpMatrix::vectorMultiply without MKL (spmatrix.cpp:346-360, 395-410);
     * complex products as libstdc++'s (a+bi)(c+di) = (ac-bd) + (ad+bc)i,
     * alpha and beta real. */
    if (!is_complex) {
        for (int r = 0; r < n; ++r) {
            double s = 0;
            for (int64_t i = rp[r]; i < rp[r + 1]; ++i) s += x[ci[i]] * val[i];
            y[r] = alpha * s + beta * y[r];
        }
        return;
    }
    for (int r = 0; r < n; ++r) {
        double sr = 0, si = 0;
        for (int64_t i = rp[r]; i < rp[r + 1]; ++i) {
            double a = x[2 * (size_t)ci[i]], b = x[2 * (size_t)ci[i] + 1];
            double c = val[2 * i], d = val[2 * i + 1];
            sr += a * c - b * d;
            si += a * d + b * c;
        }
        y[2 * r] = alpha * sr + beta * y[2 * r];
        y[2 * r + 1] = alpha * si + beta * y[2 * r + 1];
    }
}

typedef struct {
    int n; const int64_t* rp; const int32_t* ci; const double* val; int cpx;
    int M; int start; int target; const double* v; int vcpx; int doubling_all;
    double* mu;
} kpm_task;

/* inner product Re <a|b> with a possibly real (vw = 1) or complex (vw = 2) */
static double re_dot(int n, const double* a, int aw, const double* b, int bw)
{
    double s = 0;
    for (int i = 0; i < n; ++i) {
        if (aw == 1 && bw == 1) s += b[i] * a[i];
        else if (aw == 1) s += b[2 * i] * a[i];
        else if (bw == 1) s += a[2 * i] * b[i];
        else s += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
    }
    return s;
}

/* <T_j|T_j> and Re<T_next|T_j>, sequential (locality.cpp:6022-6028, 6109-6114);
 * complex terms as libstdc++'s (conj(a)*b).real() = a.re*b.re + a.im*b.im */
static void kpm_dots(int n, int w, const double* Tj, const double* Tn, double* an_an, double* anp_an)
{
    double s1 = 0, s2 = 0;
    if (w == 1) {
        for (int c = 0; c < n; ++c) { s1 += Tj[c] * Tj[c]; s2 += Tn[c] * Tj[c]; }
    } else {
        for (int c = 0; c < n; ++c) {
            s1 += Tj[2 * c] * Tj[2 * c] + Tj[2 * c + 1] * Tj[2 * c + 1];
            s2 += Tn[2 * c] * Tj[2 * c] + Tn[2 * c + 1] * Tj[2 * c + 1];
        }
    }
    *an_an = s1;
    *anp_an = s2;
}

static void kpm_one(const kpm_task* t)
{
    /* Locality::computeDosKPM (locality.cpp:5945-6136) and
     * Locality::computeDosTraceKPM (locality.cpp:6138-6346): three buffers,
     * two copies and one subtract per step, sequential dot products. */
    int n = t->n, M = t->M;
    int w = (t->cpx || (t->start == ORC_START_VECTORS && t->vcpx)) ? 2 : 1; /* vector width */
    size_t bytes = sizeof(double) * (size_t)w * (size_t)n;
    double* Tp = (double*)calloc(1, bytes + 16);
    double* Tj = (double*)calloc(1, bytes + 16);
    double* Tn = (double*)calloc(1, bytes + 16);
    double* mu = t->mu;
    int vw = t->vcpx ? 2 : 1;
    /* a real matrix with complex vectors is applied to both parts */
    int mat_cpx = t->cpx;
    double* hv = NULL;
    if (!mat_cpx && w == 2) {
        hv = (double*)calloc(2 * ((size_t)t->rp[n] + 1), sizeof(double));
        for (int64_t q = 0; q < t->rp[n]; ++q) hv[2 * q] = t->val[q];
    }
    const double* V = hv ? hv : t->val;
    int cplx = w == 2;
    if (t->start == ORC_START_TARGET) {
        Tp[(size_t)w * t->target] = 1.0;
    } else {
        for (int i = 0; i < n; ++i) {
            if (w == 1) Tp[i] = t->v[i];
            else if (vw == 1) { Tp[2 * i] = t->v[i]; Tp[2 * i + 1] = 0.0; }
            else { Tp[2 * i] = t->v[2 * i]; Tp[2 * i + 1] = t->v[2 * i + 1]; }
        }
    }
    orc_spmv(n, t->rp, t->ci, V, cplx, Tp, Tj, 1, 0);
    if (t->start == ORC_START_TARGET) {
        mu[0] = 1;
        mu[1] = Tj[(size_t)w * t->target];
    } else {
        double s = 0;
        if (vw == 1) for (int i = 0; i < n; ++i) s += t->v[i] * t->v[i];
        else for (int i = 0; i < n; ++i) s += t->v[2 * i] * t->v[2 * i] + t->v[2 * i + 1] * t->v[2 * i + 1];
        mu[0] = s;
        mu[1] = re_dot(n, t->v, vw, Tj, w);
    }
    orc_spmv(n, t->rp, t->ci, V, cplx, Tj, Tn, 2, 0);
    for (size_t c = 0; c < (size_t)w * n; ++c) Tn[c] = Tn[c] - Tp[c];
    if (t->doubling_all) {
        /* doubling identity from j = 1: mu_2 = 2<T1|T1> - mu0, mu_3 = 2<T2|T1> - mu1 */
        double an_an, anp_an;
        kpm_dots(n, w, Tj, Tn, &an_an, &anp_an);
        mu[2] = 2 * an_an - mu[0];
        mu[3] = 2 * anp_an - mu[1];
    }
    for (int j = 2; j < M / 2; ++j) {
        for (size_t c = 0; c < (size_t)w * n; ++c) { Tp[c] = Tj[c]; Tj[c] = Tn[c]; }
        if (!t->doubling_all) {
            if (t->start == ORC_START_TARGET) mu[j] = Tj[(size_t)w * t->target];
            else mu[j] = re_dot(n, t->v, vw, Tj, w);
        }
        orc_spmv(n, t->rp, t->ci, V, cplx, Tj, Tn, 2, 0);
        for (size_t c = 0; c < (size_t)w * n; ++c) Tn[c] = Tn[c] - Tp[c];
        if (t->doubling_all || j >= M / 4) {
            double an_an, anp_an;
            kpm_dots(n, w, Tj, Tn, &an_an, &anp_an);
            mu[2 * j] = 2 * an_an - mu[0];
            mu[2 * j + 1] = 2 * anp_an - mu[1];
        }
    }
    free(Tp); free(Tj); free(Tn); free(hv);
}

typedef struct { kpm_task* tasks; int ntask; int next; pthread_mutex_t lock; } kpm_pool;

static void* kpm_worker(void* arg)
{
    kpm_pool* pool = (kpm_pool*)arg;
    for (;;) {
        pthread_mutex_lock(&pool->lock);
        int id = pool->next < pool->ntask ? pool->next++ : -1;
        pthread_mutex_unlock(&pool->lock);
        if (id < 0) return NULL;
        kpm_one(&pool->tasks[id]);
    }
}

int orc_kpm(int n, const int64_t* row_ptr, const int32_t* col, const double* val, int is_complex,
            int M, int R, int start, const int32_t* targets,
            const double* vecs, int vecs_complex, int doubling_all, int nthreads, double* mu)
{
    if (M % 4 != 0 || M < 4 || R < 1 || n < 1) return -1;
    kpm_task* tasks = (kpm_task*)calloc((size_t)R, sizeof(kpm_task));
    int vw = vecs_complex ? 2 : 1;
    for (int r = 0; r < R; ++r) {
        kpm_task* t = &tasks[r];
        t->n = n; t->rp = row_ptr; t->ci = col; t->val = val; t->cpx = is_complex;
        t->M = M; t->start = start; t->target = targets ? targets[r] : 0;
        t->v = vecs ? vecs + (size_t)r * vw * n : NULL;
        t->vcpx = vecs_complex; t->doubling_all = doubling_all;
        t->mu = mu + (size_t)r * M;
        for (int j = 0; j < M; ++j) t->mu[j] = 0.0;
    }
    if (nthreads <= 1) {
        for (int r = 0; r < R; ++r) kpm_one(&tasks[r]);
    } else {
        kpm_pool pool = {tasks, R, 0, PTHREAD_MUTEX_INITIALIZER};
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, kpm_worker, &pool);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
        free(th);
    }
    free(tasks);
    return 0;
}

/* counter-based random phases: no reference counterpart (the workload's
 * random-phase start vectors).  splitmix64 finaliser on (seed, r, i). */
static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_phase_hash(uint64_t seed, uint64_t r, uint64_t i)
{
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL);
    h = mix64(h ^ (r * 0xD1B54A32D192ED03ULL));
    h = mix64(h ^ (i * 0x8CB92BA72F3D8DD7ULL));
    return h;
}

void orc_random_phase(int n, int R, uint64_t seed, int r0, double* v)
{
    for (int r = 0; r < R; ++r)
        for (int i = 0; i < n; ++i) {
            uint64_t h = orc_phase_hash(seed, (uint64_t)(r0 + r), (uint64_t)i);
            double phi = (double)(h >> 11) * (1.0 / 9007199254740992.0) * 6.283185307179586;
            v[((size_t)r * n + i) * 2] = cos(phi);
            v[((size_t)r * n + i) * 2 + 1] = sin(phi);
        }
}

/* ------------------------------------------------------------------ */
/* DoS: Param_tools::densityTransform{,Trace} (param_tools.cpp:661-768)   */

void orc_jackson(int M, double* g)
{
    for (int i = 0; i < M; ++i)
        g[i] = ((M - i) * cos((M_PI * i) / M) + sin((M_PI * i) / M) / tan(M_PI / M)) / (M);
}

void orc_energies(int M, double a, double b, double* E)
{
    for (int i = 0; i < M; ++i) E[i] = b + a * cos((i * 1.0 + 0.5) * M_PI / M);
}

/* FFTW REDFT01 (DCT-III): out_k = in_0 + 2 sum_{n>=1} in_n cos(pi n (k+1/2)/M),
 * evaluated as an O(M^2) direct sum with the angle reduced exactly in integers */
static void dct3(int M, const double* in, double* out)
{
    long long P = 4LL * M;
    for (int k = 0; k < M; ++k) {
        double s = 0.0;
        for (int q = 1; q < M; ++q) {
            long long m = ((long long)q * (2LL * k + 1)) % P;
            s += in[q] * cos(M_PI * (double)m / (2.0 * M));
        }
        out[k] = in[0] + 2.0 * s;
    }
}

void orc_dos(int M, const double* mu, double a, double b, double* rho)
{
    double* g = (double*)malloc(sizeof(double) * (size_t)M * 4);
    double *E = g + M, *in = E + M, *out = in + M;
    orc_jackson(M, g);
    orc_energies(M, a, b, E);
    for (int i = 0; i < M; ++i) in[i] = g[i] * mu[i];
    dct3(M, in, out);
    for (int i = 0; i < M; ++i) rho[i] = out[i] / (M_PI * sqrt(a * a - (E[i] - b) * (E[i] - b)));
    free(g);
}

void orc_dos_trace(int M, int S, const double* mu, double a, double b, double scale, double* rho)
{
    double* g = (double*)malloc(sizeof(double) * (size_t)M * 5);
    double *E = g + M, *in = E + M, *out = in + M, *mt = out + M;
    orc_jackson(M, g);
    orc_energies(M, a, b, E);
    for (int i = 0; i < M; ++i) {
        mt[i] = 0.0;
        for (int s = 0; s < S; ++s) mt[i] = mt[i] + mu[(size_t)s * M + i] / ((double)S);
    }
    for (int i = 0; i < M; ++i) in[i] = g[i] * mt[i];
    dct3(M, in, out);
    for (int i = 0; i < M; ++i) rho[i] = scale * out[i] / (M_PI * sqrt(a * a - (E[i] - b) * (E[i] - b)));
    free(g);
}

/* ------------------------------------------------------------------ */
/* KPM conductivity: Locality::computeCondKPM (locality.cpp:6348-6727),  */
/* real H, observable_type 1, cond_poly_par 0                            */

/* y = A x for a real CSR matrix, rows in stored order (spmatrix.cpp:346-360) */
static void csr_mv(int n, const int64_t* rp, const int32_t* ci, const double* a, const double* x, double* y)
{
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) s += a[q] * x[ci[q]];
        y[i] = s;
    }
}

int orc_cond(int n, const int64_t* rp, const int32_t* ci, const double* val, const double* vdx, int M, int T,
             const int32_t* targets, double* C)
{
    if (M < 2 || !vdx) return -1;
    double* beta = (double*)malloc(sizeof(double) * (size_t)n * M);
    double* a0 = (double*)malloc(sizeof(double) * (size_t)n * 4);
    double *a1 = a0 + n, *a2 = a1 + n, *tmp = a2 + n;
    for (int t = 0; t < T; ++t) {
        const int tg = targets[t];
        /* beta_m = T_m(H) e_t, m = 0 .. M-1 */
        for (int i = 0; i < n; ++i) beta[i] = 0.0;
        beta[tg] = 1.0;
        csr_mv(n, rp, ci, val, beta, beta + n);
        for (int m = 2; m < M; ++m) {
            double* bm = beta + (size_t)m * n;
            csr_mv(n, rp, ci, val, beta + (size_t)(m - 1) * n, bm);
            for (int i = 0; i < n; ++i) bm[i] = 2.0 * bm[i] - beta[(size_t)(m - 2) * n + i];
        }
        /* alpha_0 = -dxH e_t (locality.cpp:4326-4357) */
        for (int i = 0; i < n; ++i) tmp[i] = 0.0;
        tmp[tg] = 1.0;
        csr_mv(n, rp, ci, vdx, tmp, a0);
        for (int i = 0; i < n; ++i) a0[i] = -a0[i];
        double* row = C + (size_t)t * M * M;
        /* a_j = T_j(H) alpha_0: a_1 = H a_0, a_j = 2 H a_{j-1} - a_{j-2} */
        double *pm = a0, *pc = a1, *pn = a2;
        for (int j = 0; j < M; ++j) {
            const double* aj;
            if (j == 0) {
                aj = a0;
            } else if (j == 1) {
                csr_mv(n, rp, ci, val, a0, a1);
                aj = a1;
            } else {
                csr_mv(n, rp, ci, val, pc, pn);
                for (int i = 0; i < n; ++i) pn[i] = 2.0 * pn[i] - pm[i];
                double* r = pm;
                pm = pc;
                pc = pn;
                pn = r;
                aj = pc;
            }
            /* C[j][m] = (dxH a_j) . beta_m, summed in row order */
            csr_mv(n, rp, ci, vdx, aj, tmp);
            for (int m = 0; m < M; ++m) {
                double s = 0.0;
                const double* bm = beta + (size_t)m * n;
                for (int i = 0; i < n; ++i) s = s + tmp[i] * bm[i];
                row[(size_t)j * M + m] = s;
            }
        }
    }
    free(beta);
    free(a0);
    return 0;
}

void orc_cond_transform(int M, const double* Cin, double* out)
{
    double* g = (double*)malloc(sizeof(double) * (size_t)M * 3);
    double *col_in = g + M, *col_out = col_in + M;
    double* mid = (double*)malloc(sizeof(double) * (size_t)M * M);
    for (int i = 0; i < M; ++i)
        g[i] = ((M - i) * cos((M_PI * i) / M) + sin((M_PI * i) / M) / tan(M_PI / M)) / (sqrt(2.0) * M);
    /* rows: REDFT01 along j of 2 g_i g_j C_ij */
    for (int i = 0; i < M; ++i) {
        for (int j = 0; j < M; ++j) col_in[j] = 2 * g[i] * g[j] * Cin[(size_t)i * M + j];
        dct3(M, col_in, mid + (size_t)i * M);
    }
    /* columns: REDFT01 along i */
    for (int l = 0; l < M; ++l) {
        for (int i = 0; i < M; ++i) col_in[i] = mid[(size_t)i * M + l];
        dct3(M, col_in, col_out);
        for (int k = 0; k < M; ++k) out[(size_t)k * M + l] = col_out[k];
    }
    free(mid);
    free(g);
}
