// Matrix-free SVK tangent action y = K(x) p, FP64 or FP32, one body or a batch.
//
// The operator is the one compute_hessian_assemble_csr builds element by element and
// SolveLinearSystemCG applies through its SpMV (src/elements/FEAT10DataFunc.cuh:512-768,
// src/materials/SVK.cuh:41-67, src/solvers/LinearStaticSolver.cu:360-371), applied here
// without assembly.  The shape-function gradients are not read: per quadrature point the
// kernel reads M = J^-T (9 values) and rebuilds g_a = M dN_a/dxi from the dN/dxi table (a
// kernel argument shared by all elements).  With Ju = sum_a u_a (x) dN_a/dxi (u = x - X)
// and Jp likewise:
//     F = I + Ju M^T,  dF = Jp M^T,  E = (H + H^T + H^T H)/2 with H = F - I,
//     dE = (dF^T F + F^T dF)/2,  S = lambda tr(E) I + 2 mu E,  dS likewise from dE,
//     dP = dF S + F dS,  y_a = sum_q Q_q dN_a/dxi(q)  with  Q_q = w_q detJ_q dP M.
// Elements are processed in the force plan's 256-element blocks: the block's nodes are
// staged in LDS, element outputs are summed there, and each block writes one partial per
// node it touches; a gather kernel sums a node's partials in a fixed order.
#include <cstring>

#include "internal.h"
#include "shape.h"

namespace tlfea_hip {

template <typename T>
__device__ __forceinline__ T tan_lam(const TanMat<T>& m, int inst) {
    return m.elem_inst ? m.inst_lam[inst] : m.lam;
}
template <typename T>
__device__ __forceinline__ T tan_mu(const TanMat<T>& m, int inst) {
    return m.elem_inst ? m.inst_mu[inst] : m.mu;
}

// One quadrature point of the tangent kernel: accumulates Q_q dN_a/dxi into fa.  With
// STD the dN/dxi values are compile-time constants (q must then be a constant after
// unrolling), otherwise they come from the kernel-argument table.  MASS: 0 no mass term,
// 1 the mass of this quadrature rule (accumulated here), 2 the exact mass (the element
// volume is accumulated in vol, the term is added after the last point).  LIN: the
// positions equal the rest positions (u = 0, so H = E = S = 0 and dP = dS, the linear
// elasticity operator; the same values as the general path at u = 0, with the u terms and
// their staging left out).  AFF: M and det J are the same at every point (affine
// elements), point 0's are kept.
// MR: (lam, mu, kap) are Mooney-Rivlin's (mu10, mu01, kappa) and dP is mr_directional
// (FP64, general form only).
template <int NN, int NQ, typename T, bool STD, int MASS, bool LIN, bool AFF, bool MR = false>
__device__ __forceinline__ void tangent_qp(int q, int64_t tile, int lane, const int (&sl)[NN], const T* sU,
                                           const T* sP, int max_u, const T* __restrict__ jinv_t,
                                           const T* __restrict__ detJ_t, const DnTable<T>& dn, const tlfea_quad& quad,
                                           T lam, T mu, T kap, T eta, T lamd, T mc, T ks, T (&Mn)[9], T& dJn,
                                           T& vol, T (&fa)[NN][3]) {
    // keep the node values in LDS rather than registers across quadrature points (frees
    // ~120 VGPRs in FP64, i.e. one more wave per SIMD)
    __asm__ volatile("" ::: "memory");
    const int64_t tq = tile * NQ + q;
    T M[3][3];
#pragma unroll
    for (int c = 0; c < 9; c++) M[c / 3][c % 3] = Mn[c];
    const T dV = dJn * (T)quad.w[q];
    if constexpr (MASS == 2) vol += dV;
    // M and detJ of the next quadrature point are loaded one iteration ahead (affine
    // elements keep point 0's)
    if (q + 1 < NQ && !AFF) {
        const T* Mp = jinv_t + (tq + 1) * (9 * TILE) + lane;
#pragma unroll
        for (int c = 0; c < 9; c++) Mn[c] = Mp[c * TILE];
        dJn = detJ_t[(tq + 1) * TILE + lane];
    }
    T Ju[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Jp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    T pN[3] = {0, 0, 0};  // sum_b N_b p_b (mass term)
#pragma unroll
    for (int a = 0; a < NN; a++) {
        const int s = sl[a];
        T ua[3] = {0, 0, 0};
        if constexpr (!LIN) {
            ua[0] = sU[s];
            ua[1] = sU[max_u + s];
            ua[2] = sU[2 * max_u + s];
        }
        const T pa[3] = {sP[s], sP[max_u + s], sP[2 * max_u + s]};
        if constexpr (MASS == 1) {
            const T Na = STD ? (T)std_N<NN>(q, a) : dn.N[q][a];
#pragma unroll
            for (int i = 0; i < 3; i++) pN[i] += pa[i] * Na;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const T c = STD ? (T)std_dn<NN>(q, a, k) : dn.d[q][a][k];
            if (STD && c == T(0)) continue;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                if constexpr (!LIN) Ju[i][k] += ua[i] * c;
                Jp[i][k] += pa[i] * c;
            }
        }
    }
    T dP[3][3];
    if constexpr (LIN) {
        // dF, dE = sym(dF), dP = dS = lambda tr(dE) I + 2 mu dE
        T dF[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dF[i][j] = Jp[i][0] * M[j][0] + Jp[i][1] * M[j][1] + Jp[i][2] * M[j][2];
        T dE[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dE[i][j] = T(0.5) * (dF[i][j] + dF[j][i]);
        const T trdE = dE[0][0] + dE[1][1] + dE[2][2];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dP[i][j] = (T(2) * mu * dE[i][j] + (i == j ? lam * trdE : T(0))) * (dV * ks);
    } else if constexpr (MR) {
        static_assert(sizeof(T) == sizeof(double), "the Mooney-Rivlin tangent is FP64");
        double F[3][3], dF[3][3], dPm[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                F[i][j] = (i == j ? 1.0 : 0.0) + Ju[i][0] * M[j][0] + Ju[i][1] * M[j][1] + Ju[i][2] * M[j][2];
                dF[i][j] = Jp[i][0] * M[j][0] + Jp[i][1] * M[j][1] + Jp[i][2] * M[j][2];
            }
        mr_directional(F, dF, lam, mu, kap, dPm);
        if (eta != T(0) || lamd != T(0)) {
            // Kelvin-Voigt block: C p = F (lambda_d tr(dE) I + 2 eta dE), dE = sym(F^T dF)
            double dE[3][3], Sv[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    dE[i][j] = 0.5 * (dF[0][i] * F[0][j] + dF[1][i] * F[1][j] + dF[2][i] * F[2][j] +
                                      F[0][i] * dF[0][j] + F[1][i] * dF[1][j] + F[2][i] * dF[2][j]);
            const double trdE = dE[0][0] + dE[1][1] + dE[2][2];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) Sv[i][j] = 2.0 * eta * dE[i][j] + (i == j ? lamd * trdE : 0.0);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) dPm[i][j] += F[i][0] * Sv[0][j] + F[i][1] * Sv[1][j] + F[i][2] * Sv[2][j];
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dP[i][j] = dPm[i][j] * (dV * ks);
    } else {
        // H = F - I and dF
        T H[3][3], dF[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                H[i][j] = Ju[i][0] * M[j][0] + Ju[i][1] * M[j][1] + Ju[i][2] * M[j][2];
                dF[i][j] = Jp[i][0] * M[j][0] + Jp[i][1] * M[j][1] + Jp[i][2] * M[j][2];
            }
        T Em[3][3], dE[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const T hh = H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j];
                Em[i][j] = T(0.5) * (H[i][j] + H[j][i] + hh);
                // dF^T F + F^T dF with F = I + H
                const T s2 = dF[i][j] + dF[j][i] + dF[0][i] * H[0][j] + dF[1][i] * H[1][j] + dF[2][i] * H[2][j] +
                             H[0][i] * dF[0][j] + H[1][i] * dF[1][j] + H[2][i] * dF[2][j];
                dE[i][j] = T(0.5) * s2;
            }
        const T trE = Em[0][0] + Em[1][1] + Em[2][2], trdE = dE[0][0] + dE[1][1] + dE[2][2];
        T S[3][3], dS[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                // the Kelvin-Voigt block C p = F (lambda_d tr(dE) I + 2 eta dE) joins the
                // material term (FEAT10DataFunc.cuh:686-767 with dF = sum_b p_b (x) g_b)
                S[i][j] = T(2) * mu * Em[i][j] + (i == j ? lam * trE : T(0));
                dS[i][j] = T(2) * (mu + eta) * dE[i][j] + (i == j ? (lam + lamd) * trdE : T(0));
            }
        // dP = dF S + (I + H) dS, scaled by dV
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                dP[i][j] = (dF[i][0] * S[0][j] + dF[i][1] * S[1][j] + dF[i][2] * S[2][j] + dS[i][j] + H[i][0] * dS[0][j] +
                            H[i][1] * dS[1][j] + H[i][2] * dS[2][j]) *
                           (dV * ks);
    }
    T Q[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Q[i][k] = dP[i][0] * M[0][k] + dP[i][1] * M[1][k] + dP[i][2] * M[2][k];
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const T c = STD ? (T)std_dn<NN>(q, a, k) : dn.d[q][a][k];
            if (STD && c == T(0)) continue;
#pragma unroll
            for (int i = 0; i < 3; i++) fa[a][i] += Q[i][k] * c;
        }
    if constexpr (MASS == 1) {
        const T w = mc * dV;
#pragma unroll
        for (int a = 0; a < NN; a++) {
            const T Na = (STD ? (T)std_N<NN>(q, a) : dn.N[q][a]) * w;
#pragma unroll
            for (int i = 0; i < 3; i++) fa[a][i] += Na * pN[i];
        }
    }
}

// The rest-state operator of an affine TET10 element from its values at the four corners.
// With M and det J constant, Jp(xi) = sum_a p_a (x) dN_a/dxi and Q(xi) = C : (Jp M^T) M are
// linear in xi and y_a = det J int Q dN_a/dxi dxi has a degree-2 integrand, which the
// 5-point rule integrates exactly; so does int L_v L_w dxi = (1 + delta_vw)/120 applied to
// the linear interpolants through the corners.  At corner w (L_w = 1), with
// A = sum_c p_c (x) grad L_c over the corners and P_wj = p_w (j = w) or p of edge (w, j):
//     Jp(w) = -A + 4 sum_j P_wj (x) grad L_j,
// and with Q_w = C : (Jp(w) M^T) M det J / 120, Qs = sum_w Q_w:
//     corner a:    y_a = (4 Q_a - Qs) grad L_a,
//     edge (i, j): y = 4 [(Q_i + Qs) grad L_j + (Q_j + Qs) grad L_i].
// The same values as the 5-point evaluation to rounding, in about a third of the
// arithmetic and a third of the LDS reads; outputs go straight to the LDS accumulator.
template <typename T>
__device__ __forceinline__ T grad_l_col(const T (&R)[3][3], int d, int c) {
    // (R grad L_c)_d with grad L_0 = -(1, 1, 1), grad L_k = e_k
    return c == 0 ? -(R[d][0] + R[d][1] + R[d][2]) : R[d][c - 1];
}

// Output of node a of the element: added to the block's FP64 LDS accumulators (EM false),
// or stored element-major at sO[(a * 256 + el) * 3 + d] for the in-block node sums (EM).
template <typename T, bool EM>
__device__ __forceinline__ void vertex_out(double* sF, T* sO, int max_u, int el, int a, int slot, T v0, T v1, T v2) {
    if constexpr (EM) {
        T* o = sO + (a * PLAN_BLOCK + el) * 3;
        o[0] = v0;
        o[1] = v1;
        o[2] = v2;
    } else {
        atomicAdd(&sF[slot], (double)v0);
        atomicAdd(&sF[max_u + slot], (double)v1);
        atomicAdd(&sF[2 * max_u + slot], (double)v2);
    }
}

// pv(i, a): component i of the direction at local node a of the element.
template <typename T, bool EM, typename PV>
__device__ __forceinline__ void tangent_vertex_t10(const PV& pv, const int (&sl)[10], double* sF, T* sO, int el,
                                                   int max_u, const T (&Mn)[9], T dJ, T lam, T mu, T ks) {
    // edge node between corners w and j
    constexpr int EDGE[4][4] = {{0, 4, 6, 7}, {4, 0, 5, 8}, {6, 5, 0, 9}, {7, 8, 9, 0}};
    constexpr int EDGES[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    T M[3][3];
#pragma unroll
    for (int c = 0; c < 9; c++) M[c / 3][c % 3] = Mn[c];
    const T scale = dJ * ks * (T(1) / T(120));
    T A[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T p0 = pv(i, 0);
#pragma unroll
        for (int k = 0; k < 3; k++) A[i][k] = pv(i, k + 1) - p0;
    }
    T Q[4][3][3];
#pragma unroll
    for (int w = 0; w < 4; w++) {
        T Jp[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            T P[4];
#pragma unroll
            for (int j = 0; j < 4; j++) P[j] = pv(i, j == w ? w : EDGE[w][j]);
#pragma unroll
            for (int k = 0; k < 3; k++) Jp[i][k] = T(4) * (P[k + 1] - P[0]) - A[i][k];
        }
        T dF[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dF[i][j] = Jp[i][0] * M[j][0] + Jp[i][1] * M[j][1] + Jp[i][2] * M[j][2];
        const T ltr = lam * (dF[0][0] + dF[1][1] + dF[2][2]);
        T dP[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dP[i][j] = (mu * (dF[i][j] + dF[j][i]) + (i == j ? ltr : T(0))) * scale;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++) Q[w][i][k] = dP[i][0] * M[0][k] + dP[i][1] * M[1][k] + dP[i][2] * M[2][k];
    }
    T Qs[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Qs[i][k] = Q[0][i][k] + Q[1][i][k] + Q[2][i][k] + Q[3][i][k];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        T R[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++) R[i][k] = T(4) * Q[a][i][k] - Qs[i][k];
        vertex_out<T, EM>(sF, sO, max_u, el, a, sl[a], grad_l_col(R, 0, a), grad_l_col(R, 1, a), grad_l_col(R, 2, a));
    }
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const int ci = EDGES[m][0], cj = EDGES[m][1];
        T Ri[3][3], Rj[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                Ri[i][k] = Q[ci][i][k] + Qs[i][k];
                Rj[i][k] = Q[cj][i][k] + Qs[i][k];
            }
        T o[3];
#pragma unroll
        for (int d = 0; d < 3; d++) o[d] = T(4) * (grad_l_col(Ri, d, cj) + grad_l_col(Rj, d, ci));
        vertex_out<T, EM>(sF, sO, max_u, el, 4 + m, sl[4 + m], o[0], o[1], o[2]);
    }
}

template <int NN, int NQ, typename T, bool STD, int MASS, bool LIN, bool AFF, bool MR = false>
__global__ __launch_bounds__(PLAN_BLOCK) void k_tangent(
    int n_elem, const int64_t* __restrict__ blk_off, const int* __restrict__ unodes,
    const unsigned short* __restrict__ slot_t, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ X0, const double* __restrict__ Y0,
    const double* __restrict__ Z0, const T* __restrict__ p, const unsigned char* __restrict__ mask,
    const T* __restrict__ jinv_t, const T* __restrict__ detJ_t, DnTable<T> dn, tlfea_quad quad, TanMat<T> mat,
    T* __restrict__ part, int max_u, const int* __restrict__ flags, const int* __restrict__ blk_list,
    double* __restrict__ dot_part, const unsigned char* __restrict__ own) {
    if (flags && flags[0]) return;
    extern __shared__ unsigned char lds_raw[];
    // [3][max_u] each: accumulated output (FP64 whatever T is: LDS float atomics on nodes
    // shared by many elements of a block serialise far longer than FP64 ones on gfx950),
    // displacements (not in the rest-state form), direction (masked)
    double* sF = reinterpret_cast<double*>(lds_raw);
    T* sU = reinterpret_cast<T*>(sF + 3 * max_u);
    T* sP = LIN ? sU : sU + 3 * max_u;
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int b = blk_list ? blk_list[lb] : (int)lb;
    const int64_t u0 = blk_off[b];
    const int U = (int)(blk_off[b + 1] - u0);
    const int e = b * PLAN_BLOCK + threadIdx.x;
    int inst = 0;
    bool active = e < n_elem;
    if (active && mat.elem_inst) {
        inst = mat.elem_inst[e];
        if (mat.inst_done && mat.inst_done[inst * mat.done_stride]) active = false;
    }
    const int tile = e >> 6, lane = e & (TILE - 1);
    // slots and the first quadrature point's M / detJ are requested before the staging
    // loop so that their latency overlaps it
    int sl[NN];
    T Mn[9], dJn = T(0);
    if (active) {
#pragma unroll
        for (int a = 0; a < NN; a++) sl[a] = slot_t[((int64_t)tile * NN + a) * TILE + lane];
        const T* Mp = jinv_t + (int64_t)tile * NQ * (9 * TILE) + lane;
#pragma unroll
        for (int c = 0; c < 9; c++) Mn[c] = Mp[c * TILE];
        dJn = detJ_t[(int64_t)tile * NQ * TILE + lane];
    }
    // staging, SPT nodes per thread at a time: every node index is requested before any
    // value, so each thread waits for two dependent round trips instead of two per node.
    // With the block p.Kp shares, the masked DOFs' p^2 of the nodes this block owns is
    // summed here (sP is zero there, so the end of the kernel needs only sP . sF).
    constexpr int SPT = 4;
    double mdot = 0.0, pdot = 0.0;
    for (int base = threadIdx.x; base < U; base += SPT * PLAN_BLOCK) {
        int nid[SPT];
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            const int u = base + j * PLAN_BLOCK;
            nid[j] = u < U ? unodes[u0 + u] : -1;
        }
        T pv[SPT][3], uv[SPT][3];
        bool mk[SPT][3];
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            const int n = nid[j];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                pv[j][d] = n >= 0 ? p[3 * n + d] : T(0);
                mk[j][d] = n >= 0 && mask && mask[3 * n + d];
            }
            if constexpr (!LIN) {
                // without rest arrays x, y, z are the displacement itself (no cancellation of
                // x - X0 when u is many orders below the coordinates)
                uv[j][0] = n < 0 ? T(0) : X0 ? (T)(x[n] - X0[n]) : (T)x[n];
                uv[j][1] = n < 0 ? T(0) : X0 ? (T)(y[n] - Y0[n]) : (T)y[n];
                uv[j][2] = n < 0 ? T(0) : X0 ? (T)(z[n] - Z0[n]) : (T)z[n];
            }
        }
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            const int u = base + j * PLAN_BLOCK;
            if (u >= U) break;
            if constexpr (!LIN) {
                sU[u] = uv[j][0];
                sU[max_u + u] = uv[j][1];
                sU[2 * max_u + u] = uv[j][2];
            }
            sF[u] = sF[max_u + u] = sF[2 * max_u + u] = 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) sP[d * max_u + u] = mk[j][d] ? T(0) : pv[j][d];
            if (dot_part && own[u0 + u]) {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double pp = (double)pv[j][d] * (double)pv[j][d];
                    pdot += pp;
                    if (mk[j][d]) mdot += pp;
                }
            }
        }
    }
    __syncthreads();
    if constexpr (NN == 10 && STD && MASS == 0 && LIN && AFF) {
        if (active)
            tangent_vertex_t10<T, false>([&](int i, int a) { return sP[i * max_u + sl[a]]; }, sl, sF, nullptr, 0,
                                         max_u, Mn, dJn, tan_lam(mat, inst), tan_mu(mat, inst), mat.k_scale);
    } else if (active) {
        const T lam = MR ? mat.mu10 : tan_lam(mat, inst), mu = MR ? mat.mu01 : tan_mu(mat, inst);
        const T mc = mat.elem_inst ? (mat.inst_mc ? mat.inst_mc[inst] : T(0)) : mat.mc;
        const T ks = mat.k_scale;
        T fa[NN][3];
        T vol = T(0);
#pragma unroll
        for (int a = 0; a < NN; a++) fa[a][0] = fa[a][1] = fa[a][2] = T(0);
        if constexpr (STD) {
#pragma unroll
            for (int q = 0; q < NQ; q++)
                tangent_qp<NN, NQ, T, true, MASS, LIN, AFF, MR>(q, tile, lane, sl, sU, sP, max_u, jinv_t, detJ_t, dn,
                                                                quad, lam, mu, mat.kappa, mat.eta, mat.lam_d, mc, ks,
                                                                Mn, dJn, vol, fa);
        } else {
#pragma unroll 1
            for (int q = 0; q < NQ; q++)
                tangent_qp<NN, NQ, T, false, MASS, LIN, false, MR>(q, tile, lane, sl, sU, sP, max_u, jinv_t, detJ_t,
                                                                   dn, quad, lam, mu, mat.kappa, mat.eta, mat.lam_d, mc,
                                                                   ks, Mn, dJn, vol, fa);
        }
        if constexpr (MASS == 2) {
            // exact element mass rho V C applied to the element's p, one component at a time
            const T w = mc * vol;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                T pb[NN];
#pragma unroll
                for (int c = 0; c < NN; c++) pb[c] = sP[i * max_u + sl[c]] * w;
#pragma unroll
                for (int a = 0; a < NN; a++) {
                    T acc = T(0);
#pragma unroll
                    for (int c = 0; c < NN; c++) acc += (T)exact_mass<NN>(a, c) * pb[c];
                    fa[a][i] += acc;
                }
            }
        }
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) atomicAdd(&sF[i * max_u + sl[a]], (double)fa[a][i]);
    }
    __syncthreads();
    // this block's share of p . out, out = K p on free DOFs (sP = p) and p on masked ones
    // (sP = 0 there; their p^2 is in mdot)
    double dot = mdot;
    for (int u = threadIdx.x; u < U; u += PLAN_BLOCK) {
        T* o = part + 3 * (u0 + u);
        const T f0 = (T)sF[u], f1 = (T)sF[max_u + u], f2 = (T)sF[2 * max_u + u];
        o[0] = f0;
        o[1] = f1;
        o[2] = f2;
        if (dot_part)
            dot += (double)sP[u] * (double)f0 + (double)sP[max_u + u] * (double)f1 +
                   (double)sP[2 * max_u + u] * (double)f2;
    }
    if (dot_part) {  // uniform over the block: shares of p . out and of p . p
        __shared__ double red[2][PLAN_BLOCK / 64];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            dot += __shfl_xor(dot, o, 64);
            pdot += __shfl_xor(pdot, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            red[0][threadIdx.x >> 6] = dot;
            red[1][threadIdx.x >> 6] = pdot;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0, t = 0.0;
#pragma unroll
            for (int w = 0; w < PLAN_BLOCK / 64; w++) {
                s += red[0][w];
                t += red[1][w];
            }
            dot_part[b] = s;
            dot_part[gridDim.x + b] = t;
        }
    }
}

// The rest-state operator of an affine TET10 mesh (k_tangent's corner-value form) without
// LDS atomics: every element stores its 30 outputs element-major in LDS, and after a
// barrier each block node sums its (element, local node) pairs in the fixed order of the
// plan's gather tables (struct tlfea_force_plan), so the partials are the same bits on
// every call, and no ds_add_f64 waits (C2 FP32 pair 42.1 -> 38.0 us).  No mass term, no
// p.Kp shares, blocks of at most 1024 nodes.  LDS: the outputs (2561 x 3 of T, the last row
// zero for the tables' pads; the staged direction occupies their first 3 x max_u values
// until every element holds its 30 values in registers) and the block's tables (max_ginc
// u16): 37.5 KB on C2, four blocks per CU.
template <typename T>
__global__ __launch_bounds__(PLAN_BLOCK) void k_tangent_em(
    int n_elem, const int64_t* __restrict__ blk_off, const int* __restrict__ unodes,
    const unsigned short* __restrict__ slot_t, const T* __restrict__ p, const unsigned char* __restrict__ mask,
    const T* __restrict__ jinv_t, const T* __restrict__ detJ_t, TanMat<T> mat, T* __restrict__ part, int max_u,
    const int* __restrict__ flags, const int* __restrict__ blk_list, const int64_t* __restrict__ ginc_off,
    const unsigned short* __restrict__ ginc, const unsigned* __restrict__ gwinfo, int max_gw,
    const unsigned short* __restrict__ gperm) {
    if (flags && flags[0]) return;
    extern __shared__ unsigned char lds_raw[];
    constexpr int NO = (10 * PLAN_BLOCK + 1) * 3;
    // the staged direction shares the outputs' space: every element reads its 30 values
    // into registers before any output is stored
    T* sO = reinterpret_cast<T*>(lds_raw);
    T* sP = sO;
    unsigned short* sInc = reinterpret_cast<unsigned short*>(lds_raw + align16(sizeof(T) * NO));
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int b = blk_list ? blk_list[lb] : (int)lb;
    const int64_t u0 = blk_off[b];
    const int U = (int)(blk_off[b + 1] - u0);
    const int e = b * PLAN_BLOCK + threadIdx.x;
    int inst = 0;
    bool active = e < n_elem;
    if (active && mat.elem_inst) {
        inst = mat.elem_inst[e];
        if (mat.inst_done && mat.inst_done[inst * mat.done_stride]) active = false;
    }
    const int tile = e >> 6, lane = e & (TILE - 1);
    int sl[10];
    T Mn[9], dJn = T(0);
    if (active) {
#pragma unroll
        for (int a = 0; a < 10; a++) sl[a] = slot_t[((int64_t)tile * 10 + a) * TILE + lane];
        const T* Mp = jinv_t + (int64_t)tile * 5 * (9 * TILE) + lane;
#pragma unroll
        for (int c = 0; c < 9; c++) Mn[c] = Mp[c * TILE];
        dJn = detJ_t[(int64_t)tile * 5 * TILE + lane];
    }
    // the table rows and slots of this thread's gather positions (g = threadIdx.x + 256 r)
    const int nw = (U + 63) >> 6;
    constexpr int GR = 4;  // positions per thread: max_u <= 1024
    unsigned ginfo[GR];
    int gslot[GR];
#pragma unroll
    for (int r = 0; r < GR; r++) {
        const int g = threadIdx.x + r * PLAN_BLOCK;
        ginfo[r] = (g >> 6) < nw ? gwinfo[(int64_t)b * max_gw + (g >> 6)] : 0u;
        gslot[r] = g < U ? (int)gperm[u0 + g] : 0;
    }
    // the block's gather tables (a multiple of 64 u16 entries, 16-byte aligned)
    {
        const int64_t i0 = ginc_off[b];
        const int n16 = (int)((ginc_off[b + 1] - i0) >> 3);
        const uint4* src = reinterpret_cast<const uint4*>(ginc + i0);
        uint4* dst = reinterpret_cast<uint4*>(sInc);
        for (int k = threadIdx.x; k < n16; k += PLAN_BLOCK) dst[k] = src[k];
    }
    constexpr int SPT = 4;
    for (int base = threadIdx.x; base < U; base += SPT * PLAN_BLOCK) {
        int nid[SPT];
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            const int u = base + j * PLAN_BLOCK;
            nid[j] = u < U ? unodes[u0 + u] : -1;
        }
        T pv[SPT][3];
        bool mk[SPT][3];
#pragma unroll
        for (int j = 0; j < SPT; j++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int n = nid[j];
                pv[j][d] = n >= 0 ? p[3 * n + d] : T(0);
                mk[j][d] = n >= 0 && mask && mask[3 * n + d];
            }
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            const int u = base + j * PLAN_BLOCK;
            if (u >= U) break;
#pragma unroll
            for (int d = 0; d < 3; d++) sP[d * max_u + u] = mk[j][d] ? T(0) : pv[j][d];
        }
    }
    __syncthreads();
    T pe[10][3];
    if (active) {
#pragma unroll
        for (int a = 0; a < 10; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) pe[a][i] = sP[i * max_u + sl[a]];
    }
    __syncthreads();
    if (threadIdx.x < 3) sO[NO - 3 + threadIdx.x] = T(0);
    if (active) {
        tangent_vertex_t10<T, true>([&](int i, int a) { return pe[a][i]; }, sl, nullptr, sO, threadIdx.x, max_u, Mn,
                                    dJn, tan_lam(mat, inst), tan_mu(mat, inst), mat.k_scale);
    } else {
#pragma unroll
        for (int a = 0; a < 10; a++) vertex_out<T, true>(nullptr, sO, max_u, threadIdx.x, a, 0, T(0), T(0), T(0));
    }
    __syncthreads();
    // node sums: gather position g (wave-uniform table of position wave g >> 6)
#pragma unroll
    for (int r = 0; r < GR; r++) {
        const int g = threadIdx.x + r * PLAN_BLOCK;
        if ((g >> 6) >= nw) break;
        const unsigned info = ginfo[r];
        const unsigned short* li = sInc + (info >> 8) + (g & 63);
        const int c = (int)(info & 255u);
        T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll 4
        for (int j = 0; j < c; j++) {
            const T* o = sO + 3 * (int)li[j * 64];
            s0 += o[0];
            s1 += o[1];
            s2 += o[2];
        }
        if (g < U) {
            T* o = part + 3 * (u0 + gslot[r]);
            o[0] = s0;
            o[1] = s1;
            o[2] = s2;
        }
    }
}

// out[3i+d] = sum of node i's partials, or p for masked DOFs; optional block partials of p.out
__global__ __launch_bounds__(256) void k_gather_tangent(int n_nodes, const int* __restrict__ g_off,
                                                        const int* __restrict__ g_idx,
                                                        const double* __restrict__ part,
                                                        const unsigned char* __restrict__ mask,
                                                        const double* __restrict__ p, double* __restrict__ out,
                                                        double* __restrict__ dot_part, const int* __restrict__ flags) {
    if (flags && flags[0]) return;
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * blockDim.x + threadIdx.x;
    double dotv = 0.0;
    if (i < n_nodes) {
        double s[3] = {0.0, 0.0, 0.0};
        gather_slots3(g_idx, part, g_off[i], g_off[i + 1], s[0], s[1], s[2]);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double v = (mask && mask[3 * i + d]) ? p[3 * i + d] : s[d];
            out[3 * i + d] = v;
            dotv += p[3 * i + d] * v;
        }
    }
    if (dot_part) {
        __shared__ double red[4];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dotv += __shfl_xor(dotv, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dotv;
        __syncthreads();
        if (threadIdx.x == 0) dot_part[lb] = red[0] + red[1] + red[2] + red[3];
    }
}

// dN/dxi at the quadrature points of `quad`: TET10 corners (4 L_i - 1) grad L_i, edge k
// (i,j) of {(0,1),(1,2),(0,2),(0,3),(1,3),(2,3)}: 4 (L_i grad L_j + L_j grad L_i); TET4
// the constant gradients of the barycentric coordinates (FEAT10Data.cu:92-198).
DnTable<double> make_dn_table(int nn, const tlfea_quad& quad) {
    DnTable<double> t;
    std::memset(&t, 0, sizeof(t));
    const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int q = 0; q < quad.n && q < 5; q++) {
        const double L[4] = {1.0 - quad.x[q] - quad.y[q] - quad.z[q], quad.x[q], quad.y[q], quad.z[q]};
        if (nn == 4) {
            for (int a = 0; a < 4; a++) {
                for (int k = 0; k < 3; k++) t.d[q][a][k] = dL[a][k];
                t.N[q][a] = L[a];
            }
            continue;
        }
        for (int a = 0; a < 4; a++) t.N[q][a] = L[a] * (2.0 * L[a] - 1.0);
        for (int a = 0; a < 4; a++)
            for (int k = 0; k < 3; k++) t.d[q][a][k] = (4.0 * L[a] - 1.0) * dL[a][k];
        const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
        for (int m = 0; m < 6; m++) {
            const int i = E[m][0], j = E[m][1];
            for (int k = 0; k < 3; k++) t.d[q][4 + m][k] = 4.0 * (L[i] * dL[j][k] + L[j] * dL[i][k]);
            t.N[q][4 + m] = 4.0 * L[i] * L[j];
        }
    }
    return t;
}

template <typename T>
int launch_tangent_partials(const tlfea_force_plan* pl, const double* x, const double* y, const double* z,
                            const double* X0, const double* Y0, const double* Z0, const T* jinv_t, const T* detJ_t,
                            const tlfea_quad& quad, const TanMat<T>& mat, const unsigned char* mask, const T* p,
                            const int* flags, hipStream_t s, const int* blk_list, int n_list, bool dot) {
    const int nb = blk_list ? n_list : pl->n_blocks;
    if (dot && blk_list) {
        set_error("block p.Kp partials need every plan block");
        return TLFEA_EINVAL;
    }
    double* dot_part = dot ? pl->d_dot : nullptr;
    if (nb == 0) return TLFEA_OK;
    const DnTable<double> dn64 = make_dn_table(pl->nn, quad);
    DnTable<T> dn;
    for (int q = 0; q < 5; q++)
        for (int a = 0; a < 10; a++) {
            for (int k = 0; k < 3; k++) dn.d[q][a][k] = (T)dn64.d[q][a][k];
            dn.N[q][a] = (T)dn64.N[q][a];
        }
    // 3 FP64 arrays of max_u (output) and 6 of T (displacements, direction); 3 of T in the
    // rest-state form, which stages no displacement (more blocks per CU where LDS is the limit)
    const size_t lds9 = (3 * sizeof(double) + 6 * sizeof(T)) * (size_t)pl->max_u,
                 lds6 = (3 * sizeof(double) + 3 * sizeof(T)) * (size_t)pl->max_u;
    T* part = reinterpret_cast<T*>(pl->d_part);  // d_part holds 3*n_partials doubles
    const bool has_mass = mat.elem_inst ? mat.inst_mc != nullptr : mat.mc != T(0);
    const int mass = has_mass ? (mat.mass_exact ? 2 : 1) : 0;
    if (mat.eta != T(0) || mat.lam_d != T(0)) {
        // Kelvin-Voigt: the reference adds C_ab = (eta Fh_b (x) Fh_a + eta (g_a.g_b) F F^T
        // + lambda_d Fh_a (x) Fh_b) dV to K (FEAT10DataFunc.cuh:686-767), i.e.
        // C p = F (lambda_d tr(dE) I + 2 eta dE) per point with dE = sym(F^T dF): the SVK
        // material term with (lambda_d, eta).  Away from rest the general forms add it (to
        // dS for SVK, to A:dF for Mooney-Rivlin); at F = I it is the linear-elasticity block
        // with (lambda_d, eta), so K + C at rest is the rest-state operator with
        // lambda + lambda_d and mu + eta (after Mooney-Rivlin's equivalent Lame parameters).
        if (mat.elem_inst || has_mass) {
            set_error("Kelvin-Voigt damping in the matrix-free tangent needs one body without a mass term");
            return TLFEA_EINVAL;
        }
        if (mat.at_rest) {
            TanMat<T> eq = mat;
            if (eq.mr) {
                eq.mr = 0;
                eq.mu = T(2) * (mat.mu10 + mat.mu01);
                eq.lam = mat.kappa - T(2) * eq.mu / T(3);
            }
            eq.lam += mat.lam_d;
            eq.mu += mat.eta;
            eq.eta = eq.lam_d = T(0);
            return launch_tangent_partials<T>(pl, x, y, z, X0, Y0, Z0, jinv_t, detJ_t, quad, eq, mask, p, flags, s,
                                              blk_list, n_list, dot);
        }
        // away from rest eta and lambda_d ride into the general forms
    }
    if (mat.mr) {
        // Mooney-Rivlin: at rest its tangent is linear elasticity with mu = 2 (mu10 + mu01)
        // and lambda = kappa - 2 mu / 3 (the rest-state forms run with those); elsewhere the
        // general form with dP = A : dF (FP64, one body, standard rule, no mass term)
        if (mat.at_rest) {
            TanMat<T> eq = mat;
            eq.mr = 0;
            eq.mu = T(2) * (mat.mu10 + mat.mu01);
            eq.lam = mat.kappa - T(2) * eq.mu / T(3);
            return launch_tangent_partials<T>(pl, x, y, z, X0, Y0, Z0, jinv_t, detJ_t, quad, eq, mask, p, flags, s,
                                              blk_list, n_list, dot);
        }
        if constexpr (sizeof(T) != sizeof(double)) {
            set_error("the Mooney-Rivlin tangent is FP64 only");
            return TLFEA_EINVAL;
        } else {
            if (mass != 0 || mat.elem_inst || !is_standard_rule(pl->nn, quad)) {
                set_error("the Mooney-Rivlin tangent covers one body, no mass term, the standard rule");
                return TLFEA_EINVAL;
            }
            const size_t lds = sizeof(double) * (size_t)pl->max_u * 9;
#define TLFEA_TANGENT_MR(NN_, NQ_, AFF_)                                                                       \
    k_tangent<NN_, NQ_, double, true, 0, false, AFF_, true><<<nb, PLAN_BLOCK, lds, s>>>(                         \
        pl->n_elem, pl->d_blk_off, pl->d_unodes, pl->d_slot, x, y, z, X0, Y0, Z0, p, mask, jinv_t, detJ_t, dn, quad, \
        mat, part, pl->max_u, flags, blk_list, dot_part, pl->d_own)
            if (pl->nn == 10 && mat.affine)
                TLFEA_TANGENT_MR(10, 5, true);
            else if (pl->nn == 10)
                TLFEA_TANGENT_MR(10, 5, false);
            else
                TLFEA_TANGENT_MR(4, 1, false);
#undef TLFEA_TANGENT_MR
            TLFEA_LAUNCH_CHECK();
            return TLFEA_OK;
        }
    }
    if (pl->nn == 10) {
        // the constant-folded kernel when the rule is exactly the standard 5-point one
        bool std_rule = quad.n == 5;
        for (int q = 0; q < 5 && std_rule; q++)
            for (int a = 0; a < 10; a++)
            {
                for (int k = 0; k < 3; k++) std_rule = std_rule && dn64.d[q][a][k] == std_dn<10>(q, a, k);
                std_rule = std_rule && dn64.N[q][a] == std_N<10>(q, a);
            }
#define TLFEA_TANGENT_LAUNCH(NN_, NQ_, STD_, MASS_, LIN_, AFF_)                                              \
    k_tangent<NN_, NQ_, T, STD_, MASS_, LIN_, AFF_><<<nb, PLAN_BLOCK, LIN_ ? lds6 : lds9, s>>>(               \
        pl->n_elem, pl->d_blk_off, pl->d_unodes, pl->d_slot, x, y, z, X0, Y0, Z0, p, mask, jinv_t, detJ_t, dn, quad, \
        mat, part, pl->max_u, flags, blk_list, dot_part, pl->d_own)
#define TLFEA_TANGENT_BY_MASS(NN_, NQ_, STD_, LIN_, AFF_)        \
    do {                                                         \
        if (mass == 2)                                           \
            TLFEA_TANGENT_LAUNCH(NN_, NQ_, STD_, 2, LIN_, AFF_); \
        else if (mass == 1)                                      \
            TLFEA_TANGENT_LAUNCH(NN_, NQ_, STD_, 1, LIN_, AFF_); \
        else                                                     \
            TLFEA_TANGENT_LAUNCH(NN_, NQ_, STD_, 0, LIN_, AFF_); \
    } while (0)
        // the rest-state and affine variants have the standard rule only (elsewhere the
        // general kernel computes the same values from u = 0 and the per-point M)
        const bool aff = mat.affine != 0;
        // FP32 without a mass term or p.Kp shares: the element-major kernel (in FP64 its 61 KB
        // of outputs per block leave two blocks per CU, and the LDS-atomic kernel is faster)
        if (std_rule && mat.at_rest && aff && mass == 0 && !dot_part && pl->d_ginc && sizeof(T) == 4 &&
            pl->max_u <= 4 * PLAN_BLOCK) {
            const size_t lds = align16(sizeof(T) * (10 * PLAN_BLOCK + 1) * 3) +
                               align16(sizeof(unsigned short) * pl->max_ginc);
            k_tangent_em<T><<<nb, PLAN_BLOCK, lds, s>>>(pl->n_elem, pl->d_blk_off, pl->d_unodes, pl->d_slot, p, mask,
                                                       jinv_t, detJ_t, mat, part, pl->max_u, flags, blk_list,
                                                       pl->d_ginc_off, pl->d_ginc, pl->d_gwinfo, pl->max_gw,
                                                       pl->d_gperm);
        } else if (std_rule && mat.at_rest && aff)
            TLFEA_TANGENT_BY_MASS(10, 5, true, true, true);
        else if (std_rule && mat.at_rest)
            TLFEA_TANGENT_BY_MASS(10, 5, true, true, false);
        else if (std_rule && aff)
            TLFEA_TANGENT_BY_MASS(10, 5, true, false, true);
        else if (std_rule)
            TLFEA_TANGENT_BY_MASS(10, 5, true, false, false);
        else
            TLFEA_TANGENT_BY_MASS(10, 5, false, false, false);
    } else {
        // one quadrature point: nothing for the affine variant to skip
        if (mat.at_rest)
            TLFEA_TANGENT_BY_MASS(4, 1, true, true, false);
        else
            TLFEA_TANGENT_BY_MASS(4, 1, true, false, false);
    }
#undef TLFEA_TANGENT_BY_MASS
#undef TLFEA_TANGENT_LAUNCH
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

template int launch_tangent_partials<double>(const tlfea_force_plan*, const double*, const double*, const double*,
                                             const double*, const double*, const double*, const double*,
                                             const double*, const tlfea_quad&, const TanMat<double>&,
                                             const unsigned char*, const double*, const int*, hipStream_t,
                                             const int*, int, bool);
template int launch_tangent_partials<float>(const tlfea_force_plan*, const double*, const double*, const double*,
                                            const double*, const double*, const double*, const float*, const float*,
                                            const tlfea_quad&, const TanMat<float>&, const unsigned char*,
                                            const float*, const int*, hipStream_t, const int*, int, bool);

TanMat<double> body_tanmat(const tlfea_material& m, bool at_rest, bool affine) {
    TanMat<double> t{m.lambda, m.mu, nullptr, nullptr, nullptr, nullptr, 0};
    t.at_rest = at_rest ? 1 : 0;
    t.affine = affine ? 1 : 0;
    if (m.model == 1) {
        t.mr = 1;
        t.mu10 = m.mu10;
        t.mu01 = m.mu01;
        t.kappa = m.kappa;
    }
    t.eta = m.eta_damp;
    t.lam_d = m.lambda_damp;
    return t;
}

int launch_tangent_planned(const tlfea_force_plan* pl, const double* x, const double* y, const double* z,
                           const double* X0, const double* Y0, const double* Z0, const double* jinv_t,
                           const double* detJ_t, const tlfea_quad& quad, const tlfea_material& mat,
                           const unsigned char* mask, const double* p, double* out, double* dot_part,
                           const int* flags, hipStream_t s, bool at_rest, bool affine) {
    const TanMat<double> tm = body_tanmat(mat, at_rest, affine);
    if (int rc = launch_tangent_partials<double>(pl, x, y, z, X0, Y0, Z0, jinv_t, detJ_t, quad, tm, mask, p, flags,
                                                 s, nullptr, 0))
        return rc;
    k_gather_tangent<<<grid_for(pl->n_nodes, 256), 256, 0, s>>>(pl->n_nodes, pl->d_g_off, pl->d_g_idx, pl->d_part,
                                                                mask, p, out, dot_part, flags);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// flag[0] = 1 when some node's position differs from its rest position (every writer
// stores the same value)
__global__ void k_moved(int n, const double* __restrict__ x, const double* __restrict__ y,
                        const double* __restrict__ z, const double* __restrict__ X0, const double* __restrict__ Y0,
                        const double* __restrict__ Z0, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (x[i] != X0[i] || y[i] != Y0[i] || z[i] != Z0[i])) flag[0] = 1;
}

__global__ void k_zero_int(int* __restrict__ v) { v[0] = 0; }

int positions_at_rest(int n_nodes, const double* x, const double* y, const double* z, const double* X0,
                      const double* Y0, const double* Z0, int* d_scratch, hipStream_t s, bool* at_rest) {
    if (x == X0 && y == Y0 && z == Z0) {
        *at_rest = true;
        return TLFEA_OK;
    }
    k_zero_int<<<1, 1, 0, s>>>(d_scratch);
    k_moved<<<grid_for(n_nodes, 256), 256, 0, s>>>(n_nodes, x, y, z, X0, Y0, Z0, d_scratch);
    TLFEA_LAUNCH_CHECK();
    int h = 1;
    TLFEA_HIP_TRY(hipMemcpyAsync(&h, d_scratch, sizeof(int), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    *at_rest = h == 0;
    return TLFEA_OK;
}


// ------------------------------------------------------------------------------------
// Diagonal of the tangent, d[(a,i)] = K_(a,i),(a,i): the diagonal terms of the element
// blocks of compute_hessian_assemble_csr (FEAT10DataFunc.cuh:512-768) at the positions x,
//   dV [ (lambda + mu) (F g_a)_i^2 + |g_a|^2 (lambda tr E + mu ((F F^T)_ii - 1)) + mu |F g_a|^2 ],
// scaled by k_scale, plus c M_aa when the material carries a mass term (c dV N_a^2 per point
// for the quadrature-rule mass, c V exact_mass(a, a) for the exact one), written as block
// partials like the tangent (gathered by the same kernels).
// ------------------------------------------------------------------------------------
template <int NN, int NQ>
__global__ __launch_bounds__(PLAN_BLOCK) void k_tangent_diag(
    int n_elem, const int64_t* __restrict__ blk_off, const int* __restrict__ unodes,
    const unsigned short* __restrict__ slot_t, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ X0, const double* __restrict__ Y0,
    const double* __restrict__ Z0, const double* __restrict__ jinv_t, const double* __restrict__ detJ_t,
    DnTable<double> dn, tlfea_quad quad, TanMat<double> mat, double* __restrict__ part, int max_u) {
    extern __shared__ unsigned char lds_raw[];
    double* sU = reinterpret_cast<double*>(lds_raw);  // [3][max_u]
    double* sF = sU + 3 * max_u;                      // [3][max_u]
    const int b = blockIdx.x;
    const int64_t u0 = blk_off[b];
    const int U = (int)(blk_off[b + 1] - u0);
    for (int u = threadIdx.x; u < U; u += PLAN_BLOCK) {
        const int n = unodes[u0 + u];
        sU[u] = X0 ? x[n] - X0[n] : x[n];
        sU[max_u + u] = X0 ? y[n] - Y0[n] : y[n];
        sU[2 * max_u + u] = X0 ? z[n] - Z0[n] : z[n];
        sF[u] = sF[max_u + u] = sF[2 * max_u + u] = 0.0;
    }
    __syncthreads();
    const int e = b * PLAN_BLOCK + threadIdx.x;
    if (e < n_elem) {
        const int inst = mat.elem_inst ? mat.elem_inst[e] : 0;
        const double lam = tan_lam(mat, inst), mu = tan_mu(mat, inst);
        const double mc = mat.elem_inst ? (mat.inst_mc ? mat.inst_mc[inst] : 0.0) : mat.mc;
        const double ks = mat.k_scale;
        const int mass = mc != 0.0 ? (mat.mass_exact ? 2 : 1) : 0;
        const int tile = e >> 6, lane = e & (TILE - 1);
        int sl[NN];
#pragma unroll
        for (int a = 0; a < NN; a++) sl[a] = slot_t[((int64_t)tile * NN + a) * TILE + lane];
        double da[NN][3];
#pragma unroll
        for (int a = 0; a < NN; a++) da[a][0] = da[a][1] = da[a][2] = 0.0;
        double vol = 0.0;
#pragma unroll 1
        for (int q = 0; q < NQ; q++) {
            const int64_t tq = (int64_t)tile * NQ + q;
            double M[3][3];
#pragma unroll
            for (int c = 0; c < 9; c++) M[c / 3][c % 3] = jinv_t[(tq * 9 + c) * TILE + lane];
            const double dV = detJ_t[tq * TILE + lane] * quad.w[q];
            vol += dV;
            double Ju[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int a = 0; a < NN; a++) {
                const int s = sl[a];
                const double ua[3] = {sU[s], sU[max_u + s], sU[2 * max_u + s]};
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int i = 0; i < 3; i++) Ju[i][k] += ua[i] * dn.d[q][a][k];
            }
            double F[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    F[i][j] = (i == j ? 1.0 : 0.0) + Ju[i][0] * M[j][0] + Ju[i][1] * M[j][1] + Ju[i][2] * M[j][2];
            double trE = 0.0, FFd[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                FFd[i] = F[i][0] * F[i][0] + F[i][1] * F[i][1] + F[i][2] * F[i][2];
                // tr(F^T F) = sum of the squared entries
                trE += FFd[i];
            }
            trE = 0.5 * (trE - 3.0);
#pragma unroll
            for (int a = 0; a < NN; a++) {
                double g[3];
#pragma unroll
                for (int i = 0; i < 3; i++)
                    g[i] = M[i][0] * dn.d[q][a][0] + M[i][1] * dn.d[q][a][1] + M[i][2] * dn.d[q][a][2];
                const double gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
                double Fh[3];
#pragma unroll
                for (int i = 0; i < 3; i++) Fh[i] = F[i][0] * g[0] + F[i][1] * g[1] + F[i][2] * g[2];
                const double FhFh = Fh[0] * Fh[0] + Fh[1] * Fh[1] + Fh[2] * Fh[2];
                const double m1 = mass == 1 ? mc * dV * dn.N[q][a] * dn.N[q][a] : 0.0;
#pragma unroll
                for (int i = 0; i < 3; i++)
                    da[a][i] += ks * dV * ((lam + mu) * Fh[i] * Fh[i] + gg * (lam * trE + mu * (FFd[i] - 1.0)) +
                                           mu * FhFh) + m1;
            }
        }
        if (mass == 2)
#pragma unroll
            for (int a = 0; a < NN; a++) {
                const double m2 = mc * vol * exact_mass<NN>(a, a);
                da[a][0] += m2;
                da[a][1] += m2;
                da[a][2] += m2;
            }
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) atomicAdd(&sF[i * max_u + sl[a]], da[a][i]);
    }
    __syncthreads();
    for (int u = threadIdx.x; u < U; u += PLAN_BLOCK) {
        double* o = part + 3 * (u0 + u);
        o[0] = sF[u];
        o[1] = sF[max_u + u];
        o[2] = sF[2 * max_u + u];
    }
}

int launch_tangent_diag_partials(const tlfea_force_plan* pl, const double* x, const double* y, const double* z,
                                 const double* X0, const double* Y0, const double* Z0, const double* jinv_t,
                                 const double* detJ_t, const tlfea_quad& quad, const TanMat<double>& mat,
                                 hipStream_t s) {
    if (pl->n_blocks == 0) return TLFEA_OK;
    const DnTable<double> dn = make_dn_table(pl->nn, quad);
    const size_t lds = sizeof(double) * (size_t)pl->max_u * 6;
    if (pl->nn == 10)
        k_tangent_diag<10, 5><<<pl->n_blocks, PLAN_BLOCK, lds, s>>>(pl->n_elem, pl->d_blk_off, pl->d_unodes,
                                                                    pl->d_slot, x, y, z, X0, Y0, Z0, jinv_t, detJ_t,
                                                                    dn, quad, mat, pl->d_part, pl->max_u);
    else
        k_tangent_diag<4, 1><<<pl->n_blocks, PLAN_BLOCK, lds, s>>>(pl->n_elem, pl->d_blk_off, pl->d_unodes,
                                                                   pl->d_slot, x, y, z, X0, Y0, Z0, jinv_t, detJ_t,
                                                                   dn, quad, mat, pl->d_part, pl->max_u);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

}  // namespace tlfea_hip

using namespace tlfea_hip;

extern "C" int tlfea_tangent_apply_planned(const tlfea_force_plan* plan, const double* x, const double* y,
                                           const double* z, const double* X0, const double* Y0, const double* Z0,
                                           const double* jinv_t, const double* detJ_t, int affine,
                                           const tlfea_quad* quad,
                                           const tlfea_material* mat, const unsigned char* fixed_dof_mask,
                                           const double* p, double* out, void* stream) {
    TLFEA_REQUIRE(plan && x && y && z && X0 && Y0 && Z0 && jinv_t && detJ_t && quad && mat && p && out,
                  "null argument");
    TLFEA_REQUIRE(quad->n == (plan->nn == 10 ? 5 : 1), "quadrature size does not match the element type");
    TLFEA_REQUIRE_MODEL(*mat);
    // the linear-elasticity form when the positions are the rest positions' own arrays (the
    // only form that takes Kelvin-Voigt damping: K + C)
    const bool same = x == X0 && y == Y0 && z == Z0;
    return launch_tangent_planned(plan, x, y, z, X0, Y0, Z0, jinv_t, detJ_t, *quad, *mat, fixed_dof_mask, p, out,
                                  nullptr, nullptr, as_stream(stream), same, affine != 0);
}
