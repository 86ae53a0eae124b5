// tlfea/utils/quadrature_utils.h -- tetrahedral quadrature tables of the drop-in API
// (reference: src/utils/quadrature_utils.h:125-179): the 5-point Keast rule for TET10 and
// the centroid rule for TET4, in the reference's barycentric-column convention.
#pragma once

#include "../types.h"

namespace tlfea {
namespace Quadrature {

constexpr int N_QP_T10_5 = 5;
constexpr int N_NODE_T10_10 = 10;
constexpr int N_QP_T4_1 = 1;
constexpr int N_NODE_T4_4 = 4;

inline VectorXR make_vec(std::initializer_list<Real> v) {
    VectorXR r((std::ptrdiff_t)v.size());
    std::ptrdiff_t i = 0;
    for (Real x : v) r(i++) = x;
    return r;
}

// qp0 = centroid (weight -4/5), qp1..4 = (1/2, 1/6, 1/6, 1/6) permutations (weight 9/20),
// weights scaled by the unit-tet volume 1/6; x, y, z = barycentric coordinates 2..4.
const VectorXR tet5pt_x = make_vec({0.25, 1.0 / 6.0, 0.5, 1.0 / 6.0, 1.0 / 6.0});
const VectorXR tet5pt_y = make_vec({0.25, 1.0 / 6.0, 1.0 / 6.0, 0.5, 1.0 / 6.0});
const VectorXR tet5pt_z = make_vec({0.25, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.5});
const VectorXR tet5pt_weights = make_vec({-4.0 / 5.0 * (1.0 / 6.0), 9.0 / 20.0 * (1.0 / 6.0), 9.0 / 20.0 * (1.0 / 6.0),
                                          9.0 / 20.0 * (1.0 / 6.0), 9.0 / 20.0 * (1.0 / 6.0)});

const VectorXR tet1pt_x = make_vec({0.25});
const VectorXR tet1pt_y = make_vec({0.25});
const VectorXR tet1pt_z = make_vec({0.25});
const VectorXR tet1pt_weights = make_vec({1.0 / 6.0});

}  // namespace Quadrature
}  // namespace tlfea
