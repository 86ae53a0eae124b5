// Compile-time shape-function data of the standard element rules (TET10 with the 5-point
// rule of tlfea_quadrature, TET4 with one point), shared by the kernels that rebuild
// shape-function gradients instead of reading them.
#pragma once

namespace tlfea_hip {

// dN/dxi of node a at quadrature point q of the standard rules (TET10: the 5-point rule of
// tlfea_quadrature; TET4: constant), evaluated with the same operations as
// make_dn_table so that the folded constants equal the table's entries.
template <int NN>
__host__ __device__ constexpr double std_dn(int q, int a, int k) {
    const double dLk[4] = {-1.0, k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
    if (NN == 4) return dLk[a];
    const double s = 1.0 / 6.0;
    const double px = q == 0 ? 0.25 : (q == 2 ? 0.5 : s);
    const double py = q == 0 ? 0.25 : (q == 3 ? 0.5 : s);
    const double pz = q == 0 ? 0.25 : (q == 4 ? 0.5 : s);
    const double L[4] = {1.0 - px - py - pz, px, py, pz};
    if (a < 4) return (4.0 * L[a] - 1.0) * dLk[a];
    const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    const int i = E[a - 4][0], j = E[a - 4][1];
    return 4.0 * (L[i] * dLk[j] + L[j] * dLk[i]);
}

// Shape function values of the standard rules, as std_dn.
template <int NN>
__host__ __device__ constexpr double std_N(int q, int a) {
    if (NN == 4) return 0.25;
    const double s = 1.0 / 6.0;
    const double px = q == 0 ? 0.25 : (q == 2 ? 0.5 : s);
    const double py = q == 0 ? 0.25 : (q == 3 ? 0.5 : s);
    const double pz = q == 0 ? 0.25 : (q == 4 ? 0.5 : s);
    const double L[4] = {1.0 - px - py - pz, px, py, pz};
    if (a < 4) return L[a] * (2.0 * L[a] - 1.0);
    const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    return 4.0 * L[E[a - 4][0]] * L[E[a - 4][1]];
}

// Exactly integrated shape-function mass, int N_a N_b dV / V, of a straight-sided element
// (the integrand is degree 4 for TET10, beyond the 5-point rule).  TET10: corner-corner
// 6 or 1, corner-edge -4 when the edge holds the corner and -6 otherwise, edge-edge 32,
// 16 (edges sharing a corner) or 8 (opposite edges), all over 420; TET4: (1 + d_ab)/20.
template <int NN>
__host__ __device__ constexpr double exact_mass(int a, int b) {
    if (NN == 4) return (a == b ? 2.0 : 1.0) / 20.0;
    const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    if (a > b) {
        const int t = a;
        a = b;
        b = t;
    }
    if (b < 4) return (a == b ? 6.0 : 1.0) / 420.0;
    if (a < 4) return (E[b - 4][0] == a || E[b - 4][1] == a ? -4.0 : -6.0) / 420.0;
    if (a == b) return 32.0 / 420.0;
    const int i = E[a - 4][0], j = E[a - 4][1], k = E[b - 4][0], l = E[b - 4][1];
    return (i == k || i == l || j == k || j == l ? 16.0 : 8.0) / 420.0;
}

}  // namespace tlfea_hip
