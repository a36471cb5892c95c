// model.h -- coupling-model ids shared by host and device code.
#pragma once

namespace leo2d {

// MAT_GRAPHENE: src/materials/graphene.cpp (9-shell intralayer table,
// Fang-Kaxiras interlayer).  MAT_GRAPHENE_SIMPLE: nearest-neighbour
// t = -2.7 eV with t_perp(r) = 0.48 exp(-(r - 3.35)/0.45) eV for r < 6 A.
enum Material { MAT_GRAPHENE = 0, MAT_GRAPHENE_SIMPLE = 1 };

}  // namespace leo2d
