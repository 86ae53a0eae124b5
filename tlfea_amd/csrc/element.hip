// Element kernels for TET10 (10 nodes, 5-point Keast rule) and TET4 (4 nodes, centroid):
// shape-gradient precompute, stress, internal force (stored-P and fused), FP32 variant and
// layout converters.
//
// Work mapping: per-quadrature-point kernels run one wavefront per (tile, q) with lane =
// element within the 64-element tile, so every array in the AoSoA layout is read and
// written as one contiguous 512-byte segment per wave instruction.  Per-element kernels
// run one element per lane and read grad_t the same way (q fixed per instruction).
#include "internal.h"

namespace tlfea_hip {

// ------------------------------------------------------------------------------------
// Shape functions of the quadratic tet (corner (4L-1) dL, edge 4(L_i dL_j + L_j dL_i)),
// src/elements/FEAT10Data.cu:104-136.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void t10_dN_dxi(double xi, double eta, double zeta, double dN[10][3]) {
    const double L[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
    // dL_0 = (-1,-1,-1), dL_k = e_{k-1}
#pragma unroll
    for (int d = 0; d < 3; d++) dN[0][d] = -(4.0 * L[0] - 1.0);
#pragma unroll
    for (int k = 1; k < 4; k++)
#pragma unroll
        for (int d = 0; d < 3; d++) dN[k][d] = (d == k - 1) ? (4.0 * L[k] - 1.0) : 0.0;
    constexpr int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int i = E[k][0], j = E[k][1];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double dLi = (i == 0) ? -1.0 : ((d == i - 1) ? 1.0 : 0.0);
            const double dLj = (j == 0) ? -1.0 : ((d == j - 1) ? 1.0 : 0.0);
            dN[k + 4][d] = 4.0 * (L[i] * dLj + L[j] * dLi);
        }
    }
}

// Partial-pivot Gaussian elimination of A x = b; pivot < 1e-14 gives x = 0
// (the reference's rule, src/elements/FEAT10DataFunc.cuh:32-84).
__device__ __forceinline__ void solve3_pivot(const double A[3][3], const double b[3], double x[3]) {
    double m[3][4];
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) m[i][j] = A[i][j];
        m[i][3] = b[i];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int p = k;
        double mv = fabs(m[k][k]);
#pragma unroll
        for (int i = k + 1; i < 3; i++)
            if (fabs(m[i][k]) > mv) {
                mv = fabs(m[i][k]);
                p = i;
            }
        if (p != k) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // select-based swap keeps the arrays in registers
                const double a = m[k][j];
                double bsel = m[k][j];
#pragma unroll
                for (int r = k + 1; r < 3; r++) bsel = (r == p) ? m[r][j] : bsel;
                m[k][j] = bsel;
#pragma unroll
                for (int r = k + 1; r < 3; r++)
                    if (r == p) m[r][j] = a;
            }
        }
        if (fabs(m[k][k]) < 1e-14) {
            x[0] = x[1] = x[2] = 0.0;
            return;
        }
#pragma unroll
        for (int i = k + 1; i < 3; i++) {
            const double f = m[i][k] / m[k][k];
#pragma unroll
            for (int j = k; j < 4; j++) m[i][j] -= f * m[k][j];
        }
    }
    x[2] = m[2][3] / m[2][2];
    x[1] = (m[1][3] - m[1][2] * x[2]) / m[1][1];
    x[0] = (m[0][3] - m[0][2] * x[2] - m[0][1] * x[1]) / m[0][0];
}

// ------------------------------------------------------------------------------------
// Precompute: one lane per (element, q).
// ------------------------------------------------------------------------------------
template <int NN, int NQ>
__global__ __launch_bounds__(256) void k_precompute(int n_elem, int n_tiles, const int* __restrict__ conn,
                                                    const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, tlfea_quad quad,
                                                    double* __restrict__ grad_t, double* __restrict__ detJ_t,
                                                    double* __restrict__ jinv_t) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(gid & (TILE - 1));
    const int64_t tq = gid >> 6;  // tile*NQ + q
    if (tq >= (int64_t)n_tiles * NQ) return;
    const int q = (int)(tq % NQ);
    const int tile = (int)(tq / NQ);
    const int e = tile * TILE + lane;
    double* G = grad_t ? grad_t + tq * (NN * 3 * TILE) + lane : nullptr;
    double* Mi = jinv_t ? jinv_t + tq * (9 * TILE) + lane : nullptr;
    if (e >= n_elem) {
        if (G)
#pragma unroll
            for (int c = 0; c < NN * 3; c++) G[c * TILE] = 0.0;
        if (Mi)
#pragma unroll
            for (int c = 0; c < 9; c++) Mi[c * TILE] = 0.0;
        if (detJ_t) detJ_t[tq * TILE + lane] = 0.0;
        return;
    }
    double X[NN][3];
#pragma unroll
    for (int a = 0; a < NN; a++) {
        const int n = conn[e + (int64_t)a * n_elem];
        X[a][0] = x[n];
        X[a][1] = y[n];
        X[a][2] = z[n];
    }
    double dN[NN][3];
    if constexpr (NN == 10) {
        t10_dN_dxi(quad.x[q], quad.y[q], quad.z[q], dN);
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            dN[0][d] = -1.0;
            dN[1][d] = (d == 0) ? 1.0 : 0.0;
            dN[2][d] = (d == 1) ? 1.0 : 0.0;
            dN[3][d] = (d == 2) ? 1.0 : 0.0;
        }
    }
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) J[i][j] += X[a][i] * dN[a][j];
    if (detJ_t) detJ_t[tq * TILE + lane] = det3(J);
    double JT[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) JT[i][j] = J[j][i];
    if constexpr (NN == 10) {
        // J^T g_a = dN_a/dxi by partial pivoting (FEAT10Data.cu:180-188)
        if (G)
#pragma unroll
            for (int a = 0; a < NN; a++) {
                double g[3];
                solve3_pivot(JT, dN[a], g);
#pragma unroll
                for (int d = 0; d < 3; d++) G[(a + NN * d) * TILE] = g[d];
            }
        // M = J^-T column by column with the same pivoting (g_a = M dN_a/dxi; a pivot below
        // 1e-14 gives M = 0, hence g = 0, as above)
        if (Mi)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double ek[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
                double col[3];
                solve3_pivot(JT, ek, col);
#pragma unroll
                for (int i = 0; i < 3; i++) Mi[(3 * i + k) * TILE] = col[i];
            }
    } else {
        // explicit cofactor inverse of J^T, 1/det or 0 (FEAT4Data.cu:153-178)
        const double dt = det3(JT);
        const double id = (dt != 0.0) ? 1.0 / dt : 0.0;
        double Ti[3][3];
        Ti[0][0] = (JT[1][1] * JT[2][2] - JT[1][2] * JT[2][1]) * id;
        Ti[0][1] = -(JT[0][1] * JT[2][2] - JT[0][2] * JT[2][1]) * id;
        Ti[0][2] = (JT[0][1] * JT[1][2] - JT[0][2] * JT[1][1]) * id;
        Ti[1][0] = -(JT[1][0] * JT[2][2] - JT[1][2] * JT[2][0]) * id;
        Ti[1][1] = (JT[0][0] * JT[2][2] - JT[0][2] * JT[2][0]) * id;
        Ti[1][2] = -(JT[0][0] * JT[1][2] - JT[0][2] * JT[1][0]) * id;
        Ti[2][0] = (JT[1][0] * JT[2][1] - JT[1][1] * JT[2][0]) * id;
        Ti[2][1] = -(JT[0][0] * JT[2][1] - JT[0][1] * JT[2][0]) * id;
        Ti[2][2] = (JT[0][0] * JT[1][1] - JT[0][1] * JT[1][0]) * id;
        if (G)
#pragma unroll
            for (int a = 0; a < NN; a++)
#pragma unroll
                for (int i = 0; i < 3; i++)
                    G[(a + NN * i) * TILE] = Ti[i][0] * dN[a][0] + Ti[i][1] * dN[a][1] + Ti[i][2] * dN[a][2];
        if (Mi)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int k = 0; k < 3; k++) Mi[(3 * i + k) * TILE] = Ti[i][k];
    }
}

// ------------------------------------------------------------------------------------
// Stored stress: one lane per (element, q).
// ------------------------------------------------------------------------------------
template <int NN, int NQ>
__global__ __launch_bounds__(256) void k_calc_p(int n_elem, int n_tiles, const int* __restrict__ conn,
                                                const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z, const double* __restrict__ grad_t,
                                                tlfea_material mat, const double* __restrict__ v,
                                                double* __restrict__ F_t, double* __restrict__ P_t,
                                                double* __restrict__ Fd_t, double* __restrict__ Pv_t) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(gid & (TILE - 1));
    const int64_t tq = gid >> 6;
    if (tq >= (int64_t)n_tiles * NQ) return;
    const int tile = (int)(tq / NQ);
    const int e = tile * TILE + lane;
    const int64_t o9 = tq * 9 * TILE + lane;
    if (e >= n_elem) {
#pragma unroll
        for (int c = 0; c < 9; c++) {
            if (F_t) F_t[o9 + c * TILE] = 0.0;
            if (P_t) P_t[o9 + c * TILE] = 0.0;
            if (Fd_t) Fd_t[o9 + c * TILE] = 0.0;
            if (Pv_t) Pv_t[o9 + c * TILE] = 0.0;
        }
        return;
    }
    const double* G = grad_t + tq * (NN * 3 * TILE) + lane;
    const bool damp = (v != nullptr) && (mat.eta_damp != 0.0 || mat.lambda_damp != 0.0);
    double F[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Fd[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int a = 0; a < NN; a++) {
        const int n = conn[e + (int64_t)a * n_elem];
        const double xa[3] = {x[n], y[n], z[n]};
        const double g[3] = {G[(a)*TILE], G[(a + NN) * TILE], G[(a + 2 * NN) * TILE]};
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) F[i][j] += xa[i] * g[j];
        if (damp) {
            const double va[3] = {v[3 * n], v[3 * n + 1], v[3 * n + 2]};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) Fd[i][j] += va[i] * g[j];
        }
    }
    double P[3][3], Pv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    material_stress(F, mat, P);
    if (damp) kv_stress(F, Fd, mat.eta_damp, mat.lambda_damp, Pv);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int64_t o = o9 + (i + 3 * j) * TILE;
            if (F_t) F_t[o] = F[i][j];
            if (P_t) P_t[o] = P[i][j] + Pv[i][j];
            if (Fd_t) Fd_t[o] = Fd[i][j];
            if (Pv_t) Pv_t[o] = Pv[i][j];
        }
}

// ------------------------------------------------------------------------------------
// Internal force from stored P: one lane per element, atomics into f.
// ------------------------------------------------------------------------------------
template <int NN, int NQ>
__global__ __launch_bounds__(256) void k_fint_from_p(int n_elem, const int* __restrict__ conn,
                                                     const double* __restrict__ P_t,
                                                     const double* __restrict__ grad_t,
                                                     const double* __restrict__ detJ_t, tlfea_quad quad,
                                                     double* __restrict__ f) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    const int tile = e >> 6, lane = e & (TILE - 1);
    double fa[NN][3];
#pragma unroll
    for (int a = 0; a < NN; a++) fa[a][0] = fa[a][1] = fa[a][2] = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int64_t tq = (int64_t)tile * NQ + q;
        const double* Pp = P_t + tq * 9 * TILE + lane;
        const double* G = grad_t + tq * (NN * 3 * TILE) + lane;
        double P[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) P[i][j] = Pp[(i + 3 * j) * TILE];
        const double dV = detJ_t[tq * TILE + lane] * quad.w[q];
#pragma unroll
        for (int a = 0; a < NN; a++) {
            const double g[3] = {G[a * TILE], G[(a + NN) * TILE], G[(a + 2 * NN) * TILE]};
#pragma unroll
            for (int i = 0; i < 3; i++) fa[a][i] += (P[i][0] * g[0] + P[i][1] * g[1] + P[i][2] * g[2]) * dV;
        }
    }
#pragma unroll
    for (int a = 0; a < NN; a++) {
        const int n = conn[e + (int64_t)a * n_elem];
#pragma unroll
        for (int i = 0; i < 3; i++) unsafeAtomicAdd(&f[3 * n + i], fa[a][i]);
    }
}

// ------------------------------------------------------------------------------------
// Fused force: F -> P -> nodal forces per element in registers; one lane per element.
// ------------------------------------------------------------------------------------
template <int NN, int NQ, bool DAMP, bool MR>
__global__ __launch_bounds__(256) void k_force(int n_elem, const int* __restrict__ conn, const double* __restrict__ x,
                                               const double* __restrict__ y, const double* __restrict__ z,
                                               const double* __restrict__ grad_t, const double* __restrict__ detJ_t,
                                               tlfea_quad quad, tlfea_material mat, const double* __restrict__ v,
                                               double* __restrict__ f, const int* __restrict__ flags) {
    if (flags && (flags[0] | flags[1])) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    const int tile = e >> 6, lane = e & (TILE - 1);
    int nodes[NN];
    double X[NN][3];
    double V[DAMP ? NN : 1][3];
#pragma unroll
    for (int a = 0; a < NN; a++) {
        nodes[a] = conn[e + (int64_t)a * n_elem];
        X[a][0] = x[nodes[a]];
        X[a][1] = y[nodes[a]];
        X[a][2] = z[nodes[a]];
        if constexpr (DAMP) {
            V[a][0] = v[3 * nodes[a]];
            V[a][1] = v[3 * nodes[a] + 1];
            V[a][2] = v[3 * nodes[a] + 2];
        }
    }
    double fa[NN][3];
#pragma unroll
    for (int a = 0; a < NN; a++) fa[a][0] = fa[a][1] = fa[a][2] = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int64_t tq = (int64_t)tile * NQ + q;
        const double* G = grad_t + tq * (NN * 3 * TILE) + lane;
        double g[NN][3];
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) g[a][d] = G[(a + NN * d) * TILE];
        const double dV = detJ_t[tq * TILE + lane] * quad.w[q];
        double F[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) F[i][j] += X[a][i] * g[a][j];
        double P[3][3];
        material_stress_t<MR>(F, mat, P);
        if constexpr (DAMP) {
            double Fd[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Pv[3][3];
#pragma unroll
            for (int a = 0; a < NN; a++)
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) Fd[i][j] += V[a][i] * g[a][j];
            kv_stress(F, Fd, mat.eta_damp, mat.lambda_damp, Pv);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) P[i][j] += Pv[i][j];
        }
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int i = 0; i < 3; i++)
                fa[a][i] += (P[i][0] * g[a][0] + P[i][1] * g[a][1] + P[i][2] * g[a][2]) * dV;
    }
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) unsafeAtomicAdd(&f[3 * nodes[a] + i], fa[a][i]);
}

// FP32 fused force in displacement form: F = I + sum_a u_a (x) g_a, so the strain is not
// lost to cancellation against the identity (|x| ~ 1 m against strains ~ 1e-4).
// E = (H + H^T + H^T H)/2 with H = F - I; P = F (lambda tr(E) I + 2 mu E), which equals
// the SVK form lambda (tr(F^T F)/2 - 3/2) F + mu (F F^T F - F).
template <int NN, int NQ>
__global__ __launch_bounds__(256) void k_force_f32(int n_elem, const int* __restrict__ conn,
                                                   const float* __restrict__ ux, const float* __restrict__ uy,
                                                   const float* __restrict__ uz, const float* __restrict__ grad32_t,
                                                   const float* __restrict__ dV32_t, float lambda, float mu,
                                                   float* __restrict__ f) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    const int tile = e >> 6, lane = e & (TILE - 1);
    int nodes[NN];
    float U[NN][3];
#pragma unroll
    for (int a = 0; a < NN; a++) {
        nodes[a] = conn[e + (int64_t)a * n_elem];
        U[a][0] = ux[nodes[a]];
        U[a][1] = uy[nodes[a]];
        U[a][2] = uz[nodes[a]];
    }
    float fa[NN][3];
#pragma unroll
    for (int a = 0; a < NN; a++) fa[a][0] = fa[a][1] = fa[a][2] = 0.0f;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int64_t tq = (int64_t)tile * NQ + q;
        const float* G = grad32_t + tq * (NN * 3 * TILE) + lane;
        float g[NN][3];
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) g[a][d] = G[(a + NN * d) * TILE];
        const float dV = dV32_t[tq * TILE + lane];
        float H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) H[i][j] += U[a][i] * g[a][j];
        float Em[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                Em[i][j] = 0.5f * (H[i][j] + H[j][i] + H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j]);
        const float tr = Em[0][0] + Em[1][1] + Em[2][2];
        float S[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) S[i][j] = 2.0f * mu * Em[i][j] + (i == j ? lambda * tr : 0.0f);
        float P[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                P[i][j] = S[i][j] + H[i][0] * S[0][j] + H[i][1] * S[1][j] + H[i][2] * S[2][j];
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int i = 0; i < 3; i++)
                fa[a][i] += (P[i][0] * g[a][0] + P[i][1] * g[a][1] + P[i][2] * g[a][2]) * dV;
    }
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) unsafeAtomicAdd(&f[3 * nodes[a] + i], fa[a][i]);
}

template <int NN, int NQ>
__global__ void k_grad_to_f32(int64_t n_tq, const double* __restrict__ grad_t, const double* __restrict__ detJ_t,
                              tlfea_quad quad, float* __restrict__ g32, float* __restrict__ dV32) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tq = gid >> 6;
    if (tq >= n_tq) return;
    const int lane = (int)(gid & (TILE - 1));
    const int q = (int)(tq % NQ);
#pragma unroll
    for (int c = 0; c < NN * 3; c++) {
        const int64_t o = (tq * NN * 3 + c) * TILE + lane;
        g32[o] = (float)grad_t[o];
    }
    dV32[tq * TILE + lane] = (float)(detJ_t[tq * TILE + lane] * quad.w[q]);
}

template <int NN, int NQ>
__global__ void k_grad_to_ref(int n_elem, const double* __restrict__ grad_t, const double* __restrict__ detJ_t,
                              double* __restrict__ grad_ref, double* __restrict__ detJ_ref) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_elem * NQ) return;
    const int e = (int)(gid / NQ), q = (int)(gid % NQ);
    const int64_t tq = (int64_t)(e >> 6) * NQ + q;
    const int lane = e & (TILE - 1);
#pragma unroll
    for (int c = 0; c < NN * 3; c++) grad_ref[gid * NN * 3 + c] = grad_t[(tq * NN * 3 + c) * TILE + lane];
    if (detJ_ref) detJ_ref[gid] = detJ_t[tq * TILE + lane];
}

template <int NQ>
__global__ void k_qp_to_ref(int n_elem, const double* __restrict__ qp_t, double* __restrict__ qp_ref) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_elem * NQ) return;
    const int e = (int)(gid / NQ), q = (int)(gid % NQ);
    const int64_t tq = (int64_t)(e >> 6) * NQ + q;
    const int lane = e & (TILE - 1);
#pragma unroll
    for (int c = 0; c < 9; c++) qp_ref[gid * 9 + c] = qp_t[(tq * 9 + c) * TILE + lane];
}

}  // namespace tlfea_hip

using namespace tlfea_hip;

int tlfea_hip::launch_force(int nn, int n_elem, const int* conn, const double* x, const double* y, const double* z,
                            const double* grad_t, const double* detJ_t, const tlfea_quad& quad,
                            const tlfea_material& mat, const double* v, double* f, const int* flags, hipStream_t s) {
    if (n_elem == 0) return TLFEA_OK;
    const bool damp = (v != nullptr) && (mat.eta_damp != 0.0 || mat.lambda_damp != 0.0);
    const unsigned g = grid_for(n_elem, 256);
    const bool mr = mat.model == 1;
#define TLFEA_FORCE_ATOMIC(NN_, NQ_, D_, MR_) \
    k_force<NN_, NQ_, D_, MR_><<<g, 256, 0, s>>>(n_elem, conn, x, y, z, grad_t, detJ_t, quad, mat, D_ ? v : nullptr, f, flags)
#define TLFEA_FORCE_ATOMIC_BY(NN_, NQ_)                 \
    do {                                                \
        if (damp && mr)                                 \
            TLFEA_FORCE_ATOMIC(NN_, NQ_, true, true);   \
        else if (damp)                                  \
            TLFEA_FORCE_ATOMIC(NN_, NQ_, true, false);  \
        else if (mr)                                    \
            TLFEA_FORCE_ATOMIC(NN_, NQ_, false, true);  \
        else                                            \
            TLFEA_FORCE_ATOMIC(NN_, NQ_, false, false); \
    } while (0)
    if (nn == 10)
        TLFEA_FORCE_ATOMIC_BY(10, 5);
    else
        TLFEA_FORCE_ATOMIC_BY(4, 1);
#undef TLFEA_FORCE_ATOMIC_BY
#undef TLFEA_FORCE_ATOMIC
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// ======================================================================================
// C-ABI
// ======================================================================================
namespace {
int check_elem(int nn, int n_elem, const tlfea_quad* quad) {
    TLFEA_REQUIRE(nn == 10 || nn == 4, "unsupported element: %d nodes (10 or 4)", nn);
    TLFEA_REQUIRE(n_elem >= 0, "negative element count");
    if (quad) TLFEA_REQUIRE(quad->n == (nn == 10 ? 5 : 1), "quadrature has %d points, element needs %d", quad->n,
                            nn == 10 ? 5 : 1);
    return TLFEA_OK;
}
}  // namespace

extern "C" int tlfea_quadrature(int nn, tlfea_quad* out) {
    TLFEA_REQUIRE(out != nullptr, "null output");
    TLFEA_REQUIRE(nn == 10 || nn == 4, "unsupported element: %d nodes", nn);
    *out = tlfea_quad{};
    if (nn == 10) {
        // Keast 5-point rule (src/utils/quadrature_utils.h:134-158): centroid with weight
        // -4/5 and the four (1/2,1/6,1/6,1/6) permutations with 9/20, scaled by 1/6.
        out->n = 5;
        const double s = 1.0 / 6.0;
        const double b[5][3] = {{0.25, 0.25, 0.25}, {s, s, s}, {0.5, s, s}, {s, 0.5, s}, {s, s, 0.5}};
        for (int q = 0; q < 5; q++) {
            out->x[q] = b[q][0];
            out->y[q] = b[q][1];
            out->z[q] = b[q][2];
            out->w[q] = (q == 0 ? -4.0 / 5.0 : 9.0 / 20.0) * (1.0 / 6.0);
        }
    } else {
        out->n = 1;
        out->x[0] = out->y[0] = out->z[0] = 0.25;
        out->w[0] = 1.0 / 6.0;
    }
    return TLFEA_OK;
}

int tlfea_hip::launch_precompute(int nn, int n_elem, const int* conn, const double* x, const double* y,
                                 const double* z, const tlfea_quad& quad, double* grad_t, double* detJ_t,
                                 double* jinv_t, hipStream_t s) {
    const int n_tiles = (int)(padded(n_elem) / TILE);
    const int64_t threads = (int64_t)n_tiles * quad.n * TILE;
    if (nn == 10)
        k_precompute<10, 5><<<grid_for(threads, 256), 256, 0, s>>>(n_elem, n_tiles, conn, x, y, z, quad, grad_t,
                                                                   detJ_t, jinv_t);
    else
        k_precompute<4, 1><<<grid_for(threads, 256), 256, 0, s>>>(n_elem, n_tiles, conn, x, y, z, quad, grad_t,
                                                                  detJ_t, jinv_t);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_precompute_grad(int nn, int n_elem, const int* conn, const double* x, const double* y,
                                     const double* z, const tlfea_quad* quad, double* grad_t, double* detJ_t,
                                     void* stream) {
    if (int rc = check_elem(nn, n_elem, quad)) return rc;
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(conn && x && y && z && grad_t && detJ_t, "null pointer argument");
    return launch_precompute(nn, n_elem, conn, x, y, z, *quad, grad_t, detJ_t, nullptr, as_stream(stream));
}

extern "C" int tlfea_precompute_jinv(int nn, int n_elem, const int* conn, const double* x, const double* y,
                                     const double* z, const tlfea_quad* quad, double* jinv_t, void* stream) {
    if (int rc = check_elem(nn, n_elem, quad)) return rc;
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(conn && x && y && z && jinv_t, "null pointer argument");
    return launch_precompute(nn, n_elem, conn, x, y, z, *quad, nullptr, nullptr, jinv_t, as_stream(stream));
}

namespace tlfea_hip {
// not_affine[0] = 1 when some element's M = J^-T or det J differs between quadrature points
// by more than 1e-12 relative (straight-sided elements agree up to the rounding of
// J = sum_a X_a (x) dN_a/dxi); every writer stores the same value.
template <int NQ>
__global__ __launch_bounds__(256) void k_not_affine(int n_elem, const double* __restrict__ jinv_t,
                                                    const double* __restrict__ detJ_t, int* __restrict__ not_affine) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    const int64_t tile = e / TILE;
    const int lane = (int)(e % TILE);
    double m0[9], mx = 0.0;
    const double* M0 = jinv_t + tile * NQ * 9 * TILE + lane;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        m0[c] = M0[c * TILE];
        mx = fmax(mx, fabs(m0[c]));
    }
    const double d0 = detJ_t[tile * NQ * TILE + lane];
    bool ok = true;
#pragma unroll
    for (int q = 1; q < NQ; q++) {
        const double* Mq = jinv_t + (tile * NQ + q) * 9 * TILE + lane;
#pragma unroll
        for (int c = 0; c < 9; c++) ok = ok && fabs(Mq[c * TILE] - m0[c]) <= 1e-12 * mx;
        ok = ok && fabs(detJ_t[(tile * NQ + q) * TILE + lane] - d0) <= 1e-12 * fabs(d0);
    }
    if (!ok) not_affine[0] = 1;
}
__global__ void k_zero_flag(int* __restrict__ v) { v[0] = 0; }
}  // namespace tlfea_hip

extern "C" int tlfea_jinv_affine(int nn, int n_elem, const double* jinv_t, const double* detJ_t, int* affine,
                                 void* stream) {
    if (int rc = check_elem(nn, n_elem, nullptr)) return rc;
    TLFEA_REQUIRE(affine != nullptr, "null output");
    *affine = 1;
    if (n_elem == 0 || nn == 4) return TLFEA_OK;  // one quadrature point: nothing to compare
    TLFEA_REQUIRE(jinv_t && detJ_t, "null pointer argument");
    hipStream_t s = as_stream(stream);
    int* d_flag = nullptr;
    TLFEA_HIP_TRY(hipMalloc((void**)&d_flag, sizeof(int)));
    k_zero_flag<<<1, 1, 0, s>>>(d_flag);
    k_not_affine<5><<<grid_for(n_elem, 256), 256, 0, s>>>(n_elem, jinv_t, detJ_t, d_flag);
    int h = 1;
    const hipError_t e0 = hipGetLastError();
    const hipError_t e1 = hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(d_flag);
    if (e0 != hipSuccess || e1 != hipSuccess || e2 != hipSuccess) {
        set_error("affinity check failed");
        return TLFEA_EHIP;
    }
    *affine = h ? 0 : 1;
    return TLFEA_OK;
}

extern "C" int tlfea_calc_p(int nn, int n_elem, const int* conn, const double* x, const double* y, const double* z,
                            const double* grad_t, const tlfea_material* mat, const double* v, double* F_t, double* P_t,
                            double* Fdot_t, double* Pvis_t, void* stream) {
    if (int rc = check_elem(nn, n_elem, nullptr)) return rc;
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(conn && x && y && z && grad_t && mat, "null pointer argument");
    TLFEA_REQUIRE_MODEL(*mat);
    const int n_tiles = (int)(padded(n_elem) / TILE);
    const int nq = nn == 10 ? 5 : 1;
    const int64_t threads = (int64_t)n_tiles * nq * TILE;
    if (nn == 10)
        k_calc_p<10, 5><<<grid_for(threads, 256), 256, 0, as_stream(stream)>>>(n_elem, n_tiles, conn, x, y, z, grad_t,
                                                                               *mat, v, F_t, P_t, Fdot_t, Pvis_t);
    else
        k_calc_p<4, 1><<<grid_for(threads, 256), 256, 0, as_stream(stream)>>>(n_elem, n_tiles, conn, x, y, z, grad_t,
                                                                              *mat, v, F_t, P_t, Fdot_t, Pvis_t);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_fint_from_p(int nn, int n_elem, const int* conn, const double* P_t, const double* grad_t,
                                 const double* detJ_t, const tlfea_quad* quad, double* f, void* stream) {
    if (int rc = check_elem(nn, n_elem, quad)) return rc;
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(conn && P_t && grad_t && detJ_t && f, "null pointer argument");
    if (nn == 10)
        k_fint_from_p<10, 5><<<grid_for(n_elem, 256), 256, 0, as_stream(stream)>>>(n_elem, conn, P_t, grad_t, detJ_t,
                                                                                   *quad, f);
    else
        k_fint_from_p<4, 1><<<grid_for(n_elem, 256), 256, 0, as_stream(stream)>>>(n_elem, conn, P_t, grad_t, detJ_t,
                                                                                  *quad, f);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_internal_force(int nn, int n_elem, int n_nodes, const int* conn, const double* x,
                                    const double* y, const double* z, const double* grad_t, const double* detJ_t,
                                    const tlfea_quad* quad, const tlfea_material* mat, const double* v, double* f,
                                    int accumulate, void* stream) {
    if (int rc = check_elem(nn, n_elem, quad)) return rc;
    TLFEA_REQUIRE(n_nodes >= 0 && f && mat, "bad arguments");
    TLFEA_REQUIRE_MODEL(*mat);
    hipStream_t s = as_stream(stream);
    if (!accumulate) TLFEA_HIP_TRY(hipMemsetAsync(f, 0, sizeof(double) * 3 * (size_t)n_nodes, s));
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(conn && x && y && z && grad_t && detJ_t, "null pointer argument");
    return launch_force(nn, n_elem, conn, x, y, z, grad_t, detJ_t, *quad, *mat, v, f, nullptr, s);
}

extern "C" int tlfea_grad_to_f32(int nn, int n_elem, const double* grad_t, const double* detJ_t,
                                 const tlfea_quad* quad, float* grad32_t, float* dV32_t, void* stream) {
    if (int rc = check_elem(nn, n_elem, quad)) return rc;
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(grad_t && detJ_t && grad32_t && dV32_t, "null pointer argument");
    const int64_t n_tq = padded(n_elem) / TILE * quad->n;
    const unsigned g = grid_for(n_tq * TILE, 256);
    if (nn == 10)
        k_grad_to_f32<10, 5><<<g, 256, 0, as_stream(stream)>>>(n_tq, grad_t, detJ_t, *quad, grad32_t, dV32_t);
    else
        k_grad_to_f32<4, 1><<<g, 256, 0, as_stream(stream)>>>(n_tq, grad_t, detJ_t, *quad, grad32_t, dV32_t);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_internal_force_f32(int nn, int n_elem, int n_nodes, const int* conn, const float* ux,
                                        const float* uy, const float* uz, const float* grad32_t, const float* dV32_t,
                                        const tlfea_material* mat, float* f, int accumulate, void* stream) {
    if (int rc = check_elem(nn, n_elem, nullptr)) return rc;
    TLFEA_REQUIRE(n_nodes >= 0 && f && mat, "bad arguments");
    TLFEA_REQUIRE(mat->model == 0, "the FP32 force covers St. Venant-Kirchhoff only");
    hipStream_t s = as_stream(stream);
    if (!accumulate) TLFEA_HIP_TRY(hipMemsetAsync(f, 0, sizeof(float) * 3 * (size_t)n_nodes, s));
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(conn && ux && uy && uz && grad32_t && dV32_t, "null pointer argument");
    const unsigned g = grid_for(n_elem, 256);
    if (nn == 10)
        k_force_f32<10, 5><<<g, 256, 0, s>>>(n_elem, conn, ux, uy, uz, grad32_t, dV32_t, (float)mat->lambda,
                                             (float)mat->mu, f);
    else
        k_force_f32<4, 1><<<g, 256, 0, s>>>(n_elem, conn, ux, uy, uz, grad32_t, dV32_t, (float)mat->lambda,
                                            (float)mat->mu, f);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_grad_to_ref(int nn, int n_elem, const double* grad_t, const double* detJ_t, double* grad_ref,
                                 double* detJ_ref, void* stream) {
    if (int rc = check_elem(nn, n_elem, nullptr)) return rc;
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(grad_t && detJ_t && grad_ref, "null pointer argument");
    const int nq = nn == 10 ? 5 : 1;
    const unsigned g = grid_for((int64_t)n_elem * nq, 256);
    if (nn == 10)
        k_grad_to_ref<10, 5><<<g, 256, 0, as_stream(stream)>>>(n_elem, grad_t, detJ_t, grad_ref, detJ_ref);
    else
        k_grad_to_ref<4, 1><<<g, 256, 0, as_stream(stream)>>>(n_elem, grad_t, detJ_t, grad_ref, detJ_ref);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_qp_to_ref(int nn, int n_elem, const double* qp_t, double* qp_ref, void* stream) {
    if (int rc = check_elem(nn, n_elem, nullptr)) return rc;
    if (n_elem == 0) return TLFEA_OK;
    TLFEA_REQUIRE(qp_t && qp_ref, "null pointer argument");
    const int nq = nn == 10 ? 5 : 1;
    const unsigned g = grid_for((int64_t)n_elem * nq, 256);
    if (nn == 10)
        k_qp_to_ref<5><<<g, 256, 0, as_stream(stream)>>>(n_elem, qp_t, qp_ref);
    else
        k_qp_to_ref<1><<<g, 256, 0, as_stream(stream)>>>(n_elem, qp_t, qp_ref);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}
