// model.h -- coupling-model ids shared by host and device code.
#pragma once

namespace leo2d {

// MAT_GRAPHENE: src/materials/graphene.cpp (9-shell intralayer table,
// Fang-Kaxiras interlayer).  MAT_GRAPHENE_SIMPLE: nearest-neighbour
// t = -2.7 eV with t_perp(r) = 0.48 exp(-(r - 3.35)/0.45) eV for r < 6 A.
// MAT_GRAPHENE_SIMPLE_STRAIN: the same, with the nearest-neighbour hopping
// rescaled by the strained bond length d: t = -2.7 exp(-3.37 (d / a_cc - 1)) eV,
// a_cc = a / sqrt(3) the unstrained bond.
enum Material { MAT_GRAPHENE = 0, MAT_GRAPHENE_SIMPLE = 1, MAT_GRAPHENE_SIMPLE_STRAIN = 2 };

constexpr double kStrainBeta = 3.37;   // d ln t / d ln d of the strained nearest-neighbour hopping

}  // namespace leo2d
