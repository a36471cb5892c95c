// model.h -- coupling-model ids shared by host and device code.
#pragma once

namespace leo2d {

// MAT_GRAPHENE: src/materials/graphene.cpp (9-shell intralayer table,
// Fang-Kaxiras interlayer).  MAT_GRAPHENE_SIMPLE: nearest-neighbour
// t = -2.7 eV with t_perp(r) = 0.48 exp(-(r - 3.35)/0.45) eV for r < 6 A.
// MAT_GRAPHENE_SIMPLE_STRAIN: the same, with the nearest-neighbour hopping
// rescaled by the strained bond length d: t = -2.7 exp(-3.37 (d / a_cc - 1)) eV,
// a_cc = a / sqrt(3) the unstrained bond.
// MAT_STRAINED_GRAPHENE: src/materials/strained_graphene.{h,cpp} -- its own
// lattice and orbital constants (sqrt(3)/2 and sqrt(3)/6 to 11 digits), search
// windows of 6 (intralayer) and 8 (interlayer) cells, a four-shell intralayer
// table that becomes strain-dependent under a uniform strain or a relaxation
// field, and the graphene interlayer term.
enum Material { MAT_GRAPHENE = 0, MAT_GRAPHENE_SIMPLE = 1, MAT_GRAPHENE_SIMPLE_STRAIN = 2, MAT_STRAINED_GRAPHENE = 3 };

// Materials::intra_search_radius / inter_search_radius truncated to int
inline int intra_window(int mat) { return mat == MAT_STRAINED_GRAPHENE ? (int)(5.7 + 2.4768 / 3.) : 4; }
inline int inter_window(int mat) { return mat == MAT_STRAINED_GRAPHENE ? (int)(8. + 2.4768 / 3.) : 6; }
// Materials::lattice (materials.cpp:60-75) and orbital_pos (materials.cpp:150-160)
inline double lattice_s32(int mat) { return mat == MAT_STRAINED_GRAPHENE ? 0.86602540378 : 0.866025403784438646764; }
inline double orbital_s36(int mat) { return mat == MAT_STRAINED_GRAPHENE ? 0.28867513459 : 0.288675134594812882255; }

constexpr double kStrainBeta = 3.37;   // d ln t / d ln d of the strained nearest-neighbour hopping

}  // namespace leo2d
