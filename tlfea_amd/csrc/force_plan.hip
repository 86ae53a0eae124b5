// Block-local fused force: LDS-staged node gather and scatter, no global atomics.
//
// A workgroup of 256 lanes owns a block of up to 256 consecutive elements (one element
// per lane) and the sorted list of the U nodes they touch.  It loads those nodes'
// positions (and velocities, with damping) into LDS once, evaluates F -> P -> nodal forces
// per element with the gradients streamed from the AoSoA layout, accumulates the nodal
// forces in LDS (ds_add_f64) and writes the U block partials with coalesced plain stores.
// A node kernel sums each node's partials in block order (the in-block LDS accumulation
// order follows wave scheduling, so results agree to rounding across calls).  HBM traffic per element is
// the gradient stream plus ~3 x 8 B per (block, node) written and read back, instead of
// 30 global FP64 atomics per element.
#include <algorithm>
#include <thread>
#include <vector>

#include <cstring>

#include "internal.h"
#include "shape.h"

namespace tlfea_hip {


template <int NN, int NQ, bool DAMP, bool MR>
__global__ __launch_bounds__(PLAN_BLOCK) void k_force_planned(
    int n_elem, const int64_t* __restrict__ blk_off, const int* __restrict__ unodes,
    const unsigned short* __restrict__ slot_t, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ v, const double* __restrict__ grad_t,
    const double* __restrict__ detJ_t, tlfea_quad quad, tlfea_material mat, double* __restrict__ part, int max_u,
    const int* __restrict__ flags) {
    if (flags && (flags[0] | flags[1])) return;
    extern __shared__ double lds[];
    double* sX = lds;                 // [3][max_u]
    double* sF = lds + 3 * max_u;     // [3][max_u]
    double* sV = lds + 6 * max_u;     // [3][max_u] (DAMP only)
    const int b = (int)xcd_block(blockIdx.x, gridDim.x);
    const int64_t u0 = blk_off[b];
    const int U = (int)(blk_off[b + 1] - u0);
    for (int u = threadIdx.x; u < U; u += PLAN_BLOCK) {
        const int n = unodes[u0 + u];
        sX[u] = x[n];
        sX[max_u + u] = y[n];
        sX[2 * max_u + u] = z[n];
        sF[u] = sF[max_u + u] = sF[2 * max_u + u] = 0.0;
        if constexpr (DAMP) {
            sV[u] = v[3 * n];
            sV[max_u + u] = v[3 * n + 1];
            sV[2 * max_u + u] = v[3 * n + 2];
        }
    }
    __syncthreads();
    const int e = b * PLAN_BLOCK + threadIdx.x;
    if (e < n_elem) {
        const int tile = e >> 6, lane = e & (TILE - 1);
        int sl[NN];
#pragma unroll
        for (int a = 0; a < NN; a++) sl[a] = slot_t[((int64_t)tile * NN + a) * TILE + lane];
        double fa[NN][3];
#pragma unroll
        for (int a = 0; a < NN; a++) fa[a][0] = fa[a][1] = fa[a][2] = 0.0;
#pragma unroll 1
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
            for (int a = 0; a < NN; a++) {
                const double xa[3] = {sX[sl[a]], sX[max_u + sl[a]], sX[2 * max_u + sl[a]]};
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) F[i][j] += xa[i] * g[a][j];
            }
            double P[3][3];
            material_stress_t<MR>(F, mat, P);
            if constexpr (DAMP) {
                double Fd[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Pv[3][3];
#pragma unroll
                for (int a = 0; a < NN; a++) {
                    const double va[3] = {sV[sl[a]], sV[max_u + sl[a]], sV[2 * max_u + sl[a]]};
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) Fd[i][j] += va[i] * g[a][j];
                }
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
            for (int i = 0; i < 3; i++) atomicAdd(&sF[i * max_u + sl[a]], fa[a][i]);
    }
    __syncthreads();
    for (int u = threadIdx.x; u < U; u += PLAN_BLOCK) {
        double* o = part + 3 * (u0 + u);
        o[0] = sF[u];
        o[1] = sF[max_u + u];
        o[2] = sF[2 * max_u + u];
    }
}

// Planned force with the shape-function gradients rebuilt from the rest coordinates
// instead of read (standard rules only).  Per quadrature point q, with compile-time
// dN_a/dxi and the block's rest positions X and displacements u = x - X staged in LDS:
//     J = sum_a X_a (x) dN_a/dxi,  g_a = J^-T dN_a/dxi,  dV = det J w_q,
//     H = (sum_a u_a (x) dN_a/dxi) J^-1 = F - I,  E = (H + H^T + H^T H) / 2,
//     S = lambda tr(E) I + 2 mu E,  P = (I + H) S  (+ Kelvin-Voigt from Fdot = sum v_a (x) g_a),
//     f_a += P g_a dV = (dV P J^-T) dN_a/dxi.
// The same f_int as compute_internal_force (FEAT10DataFunc.cuh:404-460) with the gradients
// of FEAT10Data.cu:92-198: per element only 10 u16 slots are read from HBM (the gradient
// stream of k_force_planned is 1200 B per TET10 element), and E comes from the
// displacement gradient, without the cancellation of F^T F - I at small strains.
// MASS (batched implicit steps, ForceMass): 1 adds c_i M dv with the quadrature rule's
// mass (mass_matrix_qp_kernel, FEAT10Data.cu:200-271), 2 with the exactly integrated one,
// dv staged in the velocity arrays.
template <int NN, int NQ, bool DAMP, bool MR, int MASS = 0>
__global__ __launch_bounds__(PLAN_BLOCK) void k_force_rest(
    int n_elem, const int64_t* __restrict__ blk_off, const int* __restrict__ unodes,
    const unsigned short* __restrict__ slot_t, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ X0, const double* __restrict__ Y0,
    const double* __restrict__ Z0, const double* __restrict__ v, tlfea_quad quad, tlfea_material mat,
    const int* __restrict__ elem_inst, const double* __restrict__ inst_lam, const double* __restrict__ inst_mu,
    double* __restrict__ part, int max_u, const int* __restrict__ flags, int x_is_disp, ForceMass fm) {
    if (flags && (flags[0] | flags[1])) return;
    static_assert(MASS == 0 || !DAMP, "the mass term shares the velocity arrays");
    extern __shared__ double lds[];
    double* sR = lds;              // [3][max_u] rest positions
    double* sU = lds + 3 * max_u;  // [3][max_u] displacements
    double* sF = lds + 6 * max_u;  // [3][max_u] accumulated forces
    double* sV = lds + 9 * max_u;  // [3][max_u] velocities (DAMP) or dv (MASS)
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int b = fm.blk_list ? fm.blk_list[lb] : (int)lb;
    const int64_t u0 = blk_off[b];
    const int U = (int)(blk_off[b + 1] - u0);
    const int e = b * PLAN_BLOCK + threadIdx.x;
    const int tile = e >> 6, lane = e & (TILE - 1);
    bool active = e < n_elem;
    if (active && fm.inst_done && elem_inst && fm.inst_done[elem_inst[e] * fm.done_stride]) active = false;
    int sl[NN];
    if (active) {
#pragma unroll
        for (int a = 0; a < NN; a++) sl[a] = slot_t[((int64_t)tile * NN + a) * TILE + lane];
    }
    for (int u = threadIdx.x; u < U; u += PLAN_BLOCK) {
        const int n = unodes[u0 + u];
        const double X = X0[n], Y = Y0[n], Z = Z0[n];
        sR[u] = X;
        sR[max_u + u] = Y;
        sR[2 * max_u + u] = Z;
        // x, y, z are positions, or with x_is_disp the displacement itself (not read when the
        // batched Nesterov state forms the displacement below)
        if (MASS == 0 || !fm.nes_v) {
            sU[u] = x_is_disp ? x[n] : x[n] - X;
            sU[max_u + u] = x_is_disp ? y[n] : y[n] - Y;
            sU[2 * max_u + u] = x_is_disp ? z[n] : z[n] - Z;
        }
        sF[u] = sF[max_u + u] = sF[2 * max_u + u] = 0.0;
        if constexpr (DAMP) {
            sV[u] = v[3 * n];
            sV[max_u + u] = v[3 * n + 1];
            sV[2 * max_u + u] = v[3 * n + 2];
        }
        if constexpr (MASS != 0) {
            if (fm.nes_v) {
                const int inst = fm.node_inst[n];
                const double* ts = fm.nes_s + (int64_t)inst * fm.nes_sstride;
                const double beta = (ts[0] - 1.0) / ts[1];
                const bool reset = fm.nes_reset[(int64_t)inst * fm.done_stride] != 0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const int64_t k = 3 * (int64_t)n + d;
                    const double vn = fm.nes_vn[k];
                    const double vk = reset ? vn : fm.nes_v[k];
                    const double km1 = reset ? vk : fm.nes_vkm1[k];
                    const bool fx = fm.nes_mask[k] != 0;
                    const double yv = fx ? 0.0 : vk + beta * (vk - km1);
                    sV[d * max_u + u] = fx ? 0.0 : yv - vn;
                    sU[d * max_u + u] = fm.nes_un[k] + fm.dt * yv;
                }
            } else {
                const double* dvn = fm.dv + (int64_t)fm.dv_stride * n;
                sV[u] = dvn[0];
                sV[max_u + u] = dvn[1];
                sV[2 * max_u + u] = dvn[2];
            }
        }
    }
    __syncthreads();
    if (active) {
        // one material, or per instance of a batch
        const double lam = elem_inst ? inst_lam[elem_inst[e]] : mat.lambda;
        const double mu = elem_inst ? inst_mu[elem_inst[e]] : mat.mu;
        double mc = 0.0, vol = 0.0;
        if constexpr (MASS != 0) mc = elem_inst ? fm.inst_mc[elem_inst[e]] : fm.mc;
        double fa[NN][3];
#pragma unroll
        for (int a = 0; a < NN; a++) fa[a][0] = fa[a][1] = fa[a][2] = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            // node values are re-read from LDS at every quadrature point rather than held
            // in registers across them
            __asm__ volatile("" ::: "memory");
            double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Ju[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            double Jv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int a = 0; a < NN; a++) {
                const int s = sl[a];
                const double Xa[3] = {sR[s], sR[max_u + s], sR[2 * max_u + s]};
                const double ua[3] = {sU[s], sU[max_u + s], sU[2 * max_u + s]};
                double va[3] = {0, 0, 0};
                if constexpr (DAMP) {
                    va[0] = sV[s];
                    va[1] = sV[max_u + s];
                    va[2] = sV[2 * max_u + s];
                }
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const double c = std_dn<NN>(q, a, k);
                    if (c == 0.0) continue;
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        J[i][k] += Xa[i] * c;
                        Ju[i][k] += ua[i] * c;
                        if constexpr (DAMP) Jv[i][k] += va[i] * c;
                    }
                }
            }
            // J^-1 by cofactors
            const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
            const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
            const double id = 1.0 / det;
            double Ji[3][3];  // J^-1
            Ji[0][0] = c00 * id;
            Ji[1][0] = c01 * id;
            Ji[2][0] = c02 * id;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
            double H[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) H[i][j] = Ju[i][0] * Ji[0][j] + Ju[i][1] * Ji[1][j] + Ju[i][2] * Ji[2][j];
            double S[3][3];  // E first, then S in place
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    S[i][j] = 0.5 * (H[i][j] + H[j][i] + H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j]);
            const double ltrE = lam * (S[0][0] + S[1][1] + S[2][2]);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) S[i][j] = 2.0 * mu * S[i][j] + (i == j ? ltrE : 0.0);
            double P[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) P[i][j] = S[i][j] + H[i][0] * S[0][j] + H[i][1] * S[1][j] + H[i][2] * S[2][j];
            if constexpr (MR) {  // Mooney-Rivlin from F = I + H (a batch is SVK only)
                double F[3][3];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) F[i][j] = H[i][j] + (i == j ? 1.0 : 0.0);
                mr_stress(F, mat.mu10, mat.mu01, mat.kappa, P);
            }
            if constexpr (DAMP) {
                double F[3][3], Fd[3][3], Pv[3][3];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        F[i][j] = H[i][j] + (i == j ? 1.0 : 0.0);
                        Fd[i][j] = Jv[i][0] * Ji[0][j] + Jv[i][1] * Ji[1][j] + Jv[i][2] * Ji[2][j];
                    }
                kv_stress(F, Fd, mat.eta_damp, mat.lambda_damp, Pv);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) P[i][j] += Pv[i][j];
            }
            // Q = dV P J^-T, f_a += Q dN_a/dxi
            const double dV = det * quad.w[q];
            if constexpr (MASS == 2) vol += dV;
            if constexpr (MASS == 1) {
                // c dV N_a (sum_b N_b dv_b) at this point
                double Nv[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int a = 0; a < NN; a++) {
                    const double Na = std_N<NN>(q, a);
#pragma unroll
                    for (int i = 0; i < 3; i++) Nv[i] += Na * sV[i * max_u + sl[a]];
                }
#pragma unroll
                for (int a = 0; a < NN; a++) {
                    const double w = mc * dV * std_N<NN>(q, a);
#pragma unroll
                    for (int i = 0; i < 3; i++) fa[a][i] += w * Nv[i];
                }
            }
            double Q[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int k = 0; k < 3; k++)
                    Q[i][k] = (P[i][0] * Ji[k][0] + P[i][1] * Ji[k][1] + P[i][2] * Ji[k][2]) * dV;
#pragma unroll
            for (int a = 0; a < NN; a++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const double c = std_dn<NN>(q, a, k);
                    if (c == 0.0) continue;
#pragma unroll
                    for (int i = 0; i < 3; i++) fa[a][i] += Q[i][k] * c;
                }
        }
        if constexpr (MASS == 2) {
            // c V sum_b C_ab dv_b, C the exact shape-function products over V
            const double w = mc * vol;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double db[NN];
#pragma unroll
                for (int c = 0; c < NN; c++) db[c] = sV[i * max_u + sl[c]] * w;
#pragma unroll
                for (int a = 0; a < NN; a++) {
                    double acc = 0.0;
#pragma unroll
                    for (int c = 0; c < NN; c++) acc += exact_mass<NN>(a, c) * db[c];
                    fa[a][i] += acc;
                }
            }
        }
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) atomicAdd(&sF[i * max_u + sl[a]], fa[a][i]);
    }
    __syncthreads();
    for (int u = threadIdx.x; u < U; u += PLAN_BLOCK) {
        double* o = part + 3 * (u0 + u);
        o[0] = sF[u];
        o[1] = sF[max_u + u];
        o[2] = sF[2 * max_u + u];
    }
}

// f[3i+d] (+)= sum over node i's partials, in ascending partial order (deterministic).
__global__ void k_gather_partials(int n_nodes, const int* __restrict__ g_off, const int* __restrict__ g_idx,
                                  const double* __restrict__ part, double* __restrict__ f, int accumulate,
                                  const int* __restrict__ flags) {
    if (flags && (flags[0] | flags[1])) return;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    gather_slots3(g_idx, part, g_off[i], g_off[i + 1], s0, s1, s2);
    if (accumulate) {
        f[3 * i] += s0;
        f[3 * i + 1] += s1;
        f[3 * i + 2] += s2;
    } else {
        f[3 * i] = s0;
        f[3 * i + 1] = s1;
        f[3 * i + 2] = s2;
    }
}

int launch_force_planned(const tlfea_force_plan* p, const double* x, const double* y, const double* z,
                         const double* grad_t, const double* detJ_t, const tlfea_quad& quad,
                         const tlfea_material& mat, const double* v, double* f, int accumulate, const int* flags,
                         hipStream_t s) {
    const bool damp = (v != nullptr) && (mat.eta_damp != 0.0 || mat.lambda_damp != 0.0);
    const size_t lds = sizeof(double) * (size_t)p->max_u * (damp ? 9 : 6);
    const unsigned nb = (unsigned)p->n_blocks;
    if (nb > 0) {
        const bool mr = mat.model == 1;
#define TLFEA_FORCE_PLANNED(NN_, NQ_, D_, MR_)                                                                   \
    k_force_planned<NN_, NQ_, D_, MR_><<<nb, PLAN_BLOCK, lds, s>>>(p->n_elem, p->d_blk_off, p->d_unodes, p->d_slot, \
                                                                   x, y, z, D_ ? v : nullptr, grad_t, detJ_t, quad, \
                                                                   mat, p->d_part, p->max_u, flags)
#define TLFEA_FORCE_PLANNED_BY(NN_, NQ_)                  \
    do {                                                  \
        if (damp && mr)                                   \
            TLFEA_FORCE_PLANNED(NN_, NQ_, true, true);    \
        else if (damp)                                    \
            TLFEA_FORCE_PLANNED(NN_, NQ_, true, false);   \
        else if (mr)                                      \
            TLFEA_FORCE_PLANNED(NN_, NQ_, false, true);   \
        else                                              \
            TLFEA_FORCE_PLANNED(NN_, NQ_, false, false);  \
    } while (0)
        if (p->nn == 10)
            TLFEA_FORCE_PLANNED_BY(10, 5);
        else
            TLFEA_FORCE_PLANNED_BY(4, 1);
#undef TLFEA_FORCE_PLANNED_BY
#undef TLFEA_FORCE_PLANNED
        TLFEA_LAUNCH_CHECK();
    }
    k_gather_partials<<<grid_for(p->n_nodes, 256), 256, 0, s>>>(p->n_nodes, p->d_g_off, p->d_g_idx, p->d_part, f,
                                                                 accumulate, flags);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

bool is_standard_rule(int nn, const tlfea_quad& quad) {
    if (quad.n != (nn == 10 ? 5 : 1)) return false;
    if (nn == 4) return true;  // TET4 gradients are constant: any single point
    const DnTable<double> dn = make_dn_table(nn, quad);
    for (int q = 0; q < 5; q++)
        for (int a = 0; a < 10; a++)
            for (int k = 0; k < 3; k++)
                if (dn.d[q][a][k] != std_dn<10>(q, a, k)) return false;
    return true;
}

int launch_force_rest(const tlfea_force_plan* p, const double* x, const double* y, const double* z,
                      const double* X0, const double* Y0, const double* Z0, const tlfea_quad& quad,
                      const tlfea_material& mat, const double* v, double* f, int accumulate, const int* flags,
                      hipStream_t s, const int* elem_inst, const double* inst_lam, const double* inst_mu,
                      bool x_is_disp) {
    const bool damp = (v != nullptr) && (mat.eta_damp != 0.0 || mat.lambda_damp != 0.0);
    const size_t lds = sizeof(double) * (size_t)p->max_u * (damp ? 12 : 9);
    const unsigned nb = (unsigned)p->n_blocks;
    if (nb > 0) {
        const bool mr = mat.model == 1;
#define TLFEA_FORCE_REST(NN_, NQ_, D_, MR_)                                                                      \
    k_force_rest<NN_, NQ_, D_, MR_><<<nb, PLAN_BLOCK, lds, s>>>(                                                 \
        p->n_elem, p->d_blk_off, p->d_unodes, p->d_slot, x, y, z, X0, Y0, Z0, D_ ? v : nullptr, quad, mat,       \
        elem_inst, inst_lam, inst_mu, p->d_part, p->max_u, flags, x_is_disp ? 1 : 0, ForceMass{})
#define TLFEA_FORCE_REST_BY(NN_, NQ_)                 \
    do {                                              \
        if (damp && mr)                               \
            TLFEA_FORCE_REST(NN_, NQ_, true, true);   \
        else if (damp)                                \
            TLFEA_FORCE_REST(NN_, NQ_, true, false);  \
        else if (mr)                                  \
            TLFEA_FORCE_REST(NN_, NQ_, false, true);  \
        else                                          \
            TLFEA_FORCE_REST(NN_, NQ_, false, false); \
    } while (0)
        if (p->nn == 10)
            TLFEA_FORCE_REST_BY(10, 5);
        else
            TLFEA_FORCE_REST_BY(4, 1);
#undef TLFEA_FORCE_REST_BY
#undef TLFEA_FORCE_REST
        TLFEA_LAUNCH_CHECK();
    }
    k_gather_partials<<<grid_for(p->n_nodes, 256), 256, 0, s>>>(p->n_nodes, p->d_g_off, p->d_g_idx, p->d_part, f,
                                                                 accumulate, flags);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

int launch_force_rest_mass_partials(const tlfea_force_plan* p, const double* ux, const double* uy, const double* uz,
                                    const double* X0, const double* Y0, const double* Z0, const tlfea_quad& quad,
                                    const tlfea_material& mat, const int* elem_inst, const double* inst_lam,
                                    const double* inst_mu, const ForceMass& fm, hipStream_t s) {
    const bool batch = elem_inst != nullptr;
    if ((batch && (!inst_lam || !inst_mu || !fm.inst_mc)) || (!fm.dv && !fm.nes_v) || fm.dv_stride < 3 ||
        (fm.nes_v && (!batch || !fm.nes_vkm1 || !fm.nes_vn || !fm.nes_un || !fm.nes_mask || !fm.node_inst ||
                      !fm.nes_s || !fm.nes_reset))) {
        set_error("the force with a mass term needs materials, dv and mass coefficients");
        return TLFEA_EINVAL;
    }
    if (!batch && (mat.model != 0 || mat.eta_damp != 0.0 || mat.lambda_damp != 0.0)) {
        set_error("the force with a mass term covers SVK without damping");
        return TLFEA_EINVAL;
    }
    const unsigned nb = fm.blk_list ? (unsigned)fm.n_list : (unsigned)p->n_blocks;
    if (nb == 0) return TLFEA_OK;
    const size_t lds = sizeof(double) * (size_t)p->max_u * 12;
#define TLFEA_FORCE_MASS(NN_, NQ_, M_)                                                                        \
    k_force_rest<NN_, NQ_, false, false, M_><<<nb, PLAN_BLOCK, lds, s>>>(                                     \
        p->n_elem, p->d_blk_off, p->d_unodes, p->d_slot, ux, uy, uz, X0, Y0, Z0, nullptr, quad, mat, elem_inst, \
        inst_lam, inst_mu, p->d_part, p->max_u, fm.flags, 1, fm)
    if (p->nn == 10) {
        if (fm.mass_exact)
            TLFEA_FORCE_MASS(10, 5, 2);
        else
            TLFEA_FORCE_MASS(10, 5, 1);
    } else {
        if (fm.mass_exact)
            TLFEA_FORCE_MASS(4, 1, 2);
        else
            TLFEA_FORCE_MASS(4, 1, 1);
    }
#undef TLFEA_FORCE_MASS
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

__global__ void k_fixed_mask(int n_fixed, const int* __restrict__ fixed, unsigned char* __restrict__ mask) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_fixed) return;
    const int n = fixed[k];
    mask[3 * n] = mask[3 * n + 1] = mask[3 * n + 2] = 1;
}

}  // namespace tlfea_hip

using namespace tlfea_hip;

namespace {
void free_plan(tlfea_force_plan* p) {
    if (!p) return;
    (void)hipFree(p->d_blk_off);
    (void)hipFree(p->d_unodes);
    (void)hipFree(p->d_slot);
    (void)hipFree(p->d_part);
    (void)hipFree(p->d_g_off);
    (void)hipFree(p->d_g_idx);
    (void)hipFree(p->d_own);
    (void)hipFree(p->d_dot);
    (void)hipFree(p->d_gperm);
    (void)hipFree(p->d_ginc);
    (void)hipFree(p->d_ginc_off);
    (void)hipFree(p->d_gwinfo);
    delete p;
}

template <typename T>
int upload(T** d, const std::vector<T>& h) {
    *d = nullptr;
    TLFEA_HIP_TRY(hipMalloc((void**)d, sizeof(T) * (h.empty() ? 1 : h.size())));
    if (!h.empty()) TLFEA_HIP_TRY(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return TLFEA_OK;
}
}  // namespace

extern "C" int tlfea_force_plan_create(int nn, int n_elem, int n_nodes, const int* conn, tlfea_force_plan** out,
                                       void* stream) {
    TLFEA_REQUIRE(out != nullptr, "null output handle");
    *out = nullptr;
    TLFEA_REQUIRE(nn == 10 || nn == 4, "unsupported element: %d nodes", nn);
    TLFEA_REQUIRE(n_elem > 0 && n_nodes > 0 && conn, "empty mesh or null connectivity");
    hipStream_t s = as_stream(stream);
    std::vector<int> h_conn((size_t)n_elem * nn);
    TLFEA_HIP_TRY(hipMemcpyAsync(h_conn.data(), conn, sizeof(int) * h_conn.size(), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    for (int v : h_conn)
        if (v < 0 || v >= n_nodes) {
            set_error("connectivity entry %d outside [0, %d)", v, n_nodes);
            return TLFEA_EINVAL;
        }
    auto* P = new tlfea_force_plan();
    P->nn = nn;
    P->n_elem = n_elem;
    P->n_nodes = n_nodes;
    P->block_elems = PLAN_BLOCK;
    const int nb = (n_elem + PLAN_BLOCK - 1) / PLAN_BLOCK;
    P->n_blocks = nb;
    const int64_t n_tiles = padded(n_elem) / TILE;
    // blocks are independent: host threads take contiguous ranges of them, and the results
    // are joined in block order (the plan does not depend on the thread count)
    const int nthreads = std::max(1, std::min<int>(nb, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    auto parallel_blocks = [&](auto&& fn) {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; t++)
            pool.emplace_back([&, t] {
                for (int b = (int)((int64_t)nb * t / nthreads); b < (int)((int64_t)nb * (t + 1) / nthreads); b++) fn(b);
            });
        for (auto& th : pool) th.join();
    };
    std::vector<int64_t> blk_off(nb + 1, 0);
    std::vector<std::vector<int>> blk_nodes(nb);
    std::vector<unsigned short> slot((size_t)n_tiles * nn * TILE, 0);
    parallel_blocks([&](int b) {
        const int e0 = b * PLAN_BLOCK, e1 = std::min(n_elem, e0 + PLAN_BLOCK);
        std::vector<int>& buf = blk_nodes[b];
        buf.reserve((size_t)(e1 - e0) * nn);
        for (int e = e0; e < e1; e++)
            for (int a = 0; a < nn; a++) buf.push_back(h_conn[e + (size_t)a * n_elem]);
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        for (int e = e0; e < e1; e++)
            for (int a = 0; a < nn; a++) {
                const int node = h_conn[e + (size_t)a * n_elem];
                const int sidx = (int)(std::lower_bound(buf.begin(), buf.end(), node) - buf.begin());
                slot[((size_t)(e >> 6) * nn + a) * TILE + (e & (TILE - 1))] = (unsigned short)sidx;
            }
    });
    int max_u = 0;
    for (int b = 0; b < nb; b++) {
        const int U = (int)blk_nodes[b].size();
        max_u = std::max(max_u, U);
        blk_off[b + 1] = blk_off[b] + U;
    }
    std::vector<int> unodes((size_t)blk_off[nb]);
    parallel_blocks([&](int b) {
        std::copy(blk_nodes[b].begin(), blk_nodes[b].end(), unodes.begin() + blk_off[b]);
        std::vector<int>().swap(blk_nodes[b]);
    });
    P->max_u = max_u;
    P->n_partials = blk_off[nb];
    // LDS: up to 9 arrays of max_u doubles (positions, forces, velocities) within the
    // 160 KiB of a gfx950 CU
    if ((size_t)max_u * 9 * sizeof(double) > 160 * 1024 || max_u > 65535) {
        free_plan(P);
        set_error("a block of %d elements touches %d nodes: exceeds the LDS budget", PLAN_BLOCK, max_u);
        return TLFEA_ECAPACITY;
    }
    // node -> ascending partial positions
    std::vector<int> g_off(n_nodes + 1, 0);
    for (int u : unodes) g_off[u + 1]++;
    for (int i = 0; i < n_nodes; i++) g_off[i + 1] += g_off[i];
    if (unodes.size() >= (size_t)INT32_MAX) {
        free_plan(P);
        set_error("%zu block partials: beyond the plan's 32-bit partial index", unodes.size());
        return TLFEA_ECAPACITY;
    }
    std::vector<int> g_idx(unodes.size());
    std::vector<int> cur(g_off.begin(), g_off.end() - 1);
    for (int64_t k = 0; k < (int64_t)unodes.size(); k++) g_idx[cur[unodes[k]]++] = (int)k;
    // a node's first partial in block order: where the fused CG counts its masked DOFs
    std::vector<unsigned char> own(unodes.size(), 0), seen(n_nodes, 0);
    for (size_t k = 0; k < unodes.size(); k++)
        if (!seen[unodes[k]]) {
            seen[unodes[k]] = 1;
            own[k] = 1;
        }
    std::vector<double> dot0(2 * (size_t)nb, 0.0);
    // TET10: the tables of the in-block node sums (struct tlfea_force_plan).  Positions are
    // sorted by their pair count, largest first, so that the 64 lanes of a wave loop over
    // about the same number of pairs.
    std::vector<unsigned short> gperm, ginc;
    std::vector<int64_t> ginc_off(nb + 1, 0);
    std::vector<unsigned> gwinfo;
    // a node touched by 256 or more pairs of one block (not a tetrahedral mesh one meets)
    // leaves the plan without tables: the kernels then sum with LDS atomics
    bool tables_ok = nn == 10;
    if (tables_ok) {
        const int max_gw = (max_u + 63) / 64;
        gperm.assign(unodes.size(), 0);
        gwinfo.assign((size_t)nb * max_gw, 0);
        auto slot_of = [&](int e, int a) { return (int)slot[((size_t)(e >> 6) * nn + a) * TILE + (e & (TILE - 1))]; };
        // each block's table built by its thread (offsets within the block), then joined
        std::vector<std::vector<unsigned short>> blk_tab(nb);
        std::vector<unsigned char> blk_ok(nb, 1);
        parallel_blocks([&](int b) {
            std::vector<int> cnt, start, fill, pos;
            std::vector<unsigned short> lst;
            std::vector<unsigned short>& tab = blk_tab[b];
            const int e0 = b * PLAN_BLOCK, e1 = std::min(n_elem, e0 + PLAN_BLOCK);
            const int U = (int)(blk_off[b + 1] - blk_off[b]);
            cnt.assign(U, 0);
            for (int e = e0; e < e1; e++)
                for (int a = 0; a < nn; a++) cnt[slot_of(e, a)]++;
            start.assign(U + 1, 0);
            for (int u = 0; u < U; u++) start[u + 1] = start[u] + cnt[u];
            lst.resize(start[U]);
            fill.assign(start.begin(), start.end() - 1);
            for (int e = e0; e < e1; e++)
                for (int a = 0; a < nn; a++) lst[fill[slot_of(e, a)]++] = (unsigned short)(a * PLAN_BLOCK + (e - e0));
            pos.resize(U);
            for (int u = 0; u < U; u++) pos[u] = u;
            std::stable_sort(pos.begin(), pos.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
            for (int w = 0; w * 64 < U; w++) {
                int c = 0;
                for (int l = 0; l < 64 && w * 64 + l < U; l++) c = std::max(c, cnt[pos[w * 64 + l]]);
                if (c >= 256) blk_ok[b] = 0;
                gwinfo[(size_t)b * max_gw + w] = (unsigned)(tab.size() << 8) | (unsigned)(c & 255);
                for (int j = 0; j < c; j++)
                    for (int l = 0; l < 64; l++) {
                        const int g = w * 64 + l;
                        tab.push_back(g < U && j < cnt[pos[g]] ? lst[start[pos[g]] + j]
                                                               : (unsigned short)(10 * PLAN_BLOCK));
                    }
            }
            for (int g = 0; g < U; g++) gperm[blk_off[b] + g] = (unsigned short)pos[g];
        });
        int max_ginc = 0;
        for (int b = 0; b < nb; b++) {
            tables_ok = tables_ok && blk_ok[b];
            ginc_off[b + 1] = ginc_off[b] + (int64_t)blk_tab[b].size();
            max_ginc = std::max(max_ginc, (int)blk_tab[b].size());
        }
        if (tables_ok) {
            ginc.resize((size_t)ginc_off[nb]);
            parallel_blocks([&](int b) { std::copy(blk_tab[b].begin(), blk_tab[b].end(), ginc.begin() + ginc_off[b]); });
        }
        P->max_gw = max_gw;
        P->max_ginc = max_ginc;
    }
    int rc = TLFEA_OK;
    if ((rc = upload(&P->d_blk_off, blk_off)) || (rc = upload(&P->d_unodes, unodes)) ||
        (rc = upload(&P->d_slot, slot)) || (rc = upload(&P->d_g_off, g_off)) || (rc = upload(&P->d_g_idx, g_idx)) ||
        (rc = upload(&P->d_own, own)) || (rc = upload(&P->d_dot, dot0))) {
        free_plan(P);
        return rc;
    }
    if (tables_ok && ((rc = upload(&P->d_gperm, gperm)) || (rc = upload(&P->d_ginc, ginc)) ||
                     (rc = upload(&P->d_ginc_off, ginc_off)) || (rc = upload(&P->d_gwinfo, gwinfo)))) {
        free_plan(P);
        return rc;
    }
    if (hipMalloc((void**)&P->d_part, sizeof(double) * 3 * (size_t)std::max<int64_t>(1, P->n_partials)) !=
        hipSuccess) {
        free_plan(P);
        set_error("allocation of the block partials failed");
        return TLFEA_EHIP;
    }
    *out = P;
    return TLFEA_OK;
}

extern "C" int tlfea_force_plan_destroy(tlfea_force_plan* p) {
    free_plan(p);
    return TLFEA_OK;
}

extern "C" int tlfea_force_plan_sizes(const tlfea_force_plan* p, int* block_elems, int* n_blocks,
                                      int64_t* n_partials, int* max_block_nodes) {
    TLFEA_REQUIRE(p != nullptr, "null plan");
    if (block_elems) *block_elems = p->block_elems;
    if (n_blocks) *n_blocks = p->n_blocks;
    if (n_partials) *n_partials = p->n_partials;
    if (max_block_nodes) *max_block_nodes = p->max_u;
    return TLFEA_OK;
}

extern "C" int tlfea_internal_force_planned(const tlfea_force_plan* plan, const double* x, const double* y,
                                            const double* z, const double* grad_t, const double* detJ_t,
                                            const tlfea_quad* quad, const tlfea_material* mat, const double* v,
                                            double* f, int accumulate, void* stream) {
    TLFEA_REQUIRE(plan && x && y && z && grad_t && detJ_t && quad && mat && f, "null argument");
    TLFEA_REQUIRE_MODEL(*mat);
    TLFEA_REQUIRE(quad->n == (plan->nn == 10 ? 5 : 1), "quadrature does not match the element");
    return launch_force_planned(plan, x, y, z, grad_t, detJ_t, *quad, *mat, v, f, accumulate, nullptr,
                                as_stream(stream));
}

extern "C" int tlfea_internal_force_rest(const tlfea_force_plan* plan, const double* x, const double* y,
                                         const double* z, const double* X0, const double* Y0, const double* Z0,
                                         const tlfea_quad* quad, const tlfea_material* mat, const double* v,
                                         double* f, int accumulate, void* stream) {
    TLFEA_REQUIRE(plan && x && y && z && X0 && Y0 && Z0 && quad && mat && f, "null argument");
    TLFEA_REQUIRE_MODEL(*mat);
    TLFEA_REQUIRE(is_standard_rule(plan->nn, *quad),
                  "the gradients are rebuilt for the standard rule only (TET10: 5 points, TET4: 1 point)");
    const bool damp = (v != nullptr) && (mat->eta_damp != 0.0 || mat->lambda_damp != 0.0);
    if ((size_t)plan->max_u * (damp ? 12 : 9) * sizeof(double) > 160 * 1024) {
        set_error("a block touches %d nodes: beyond the LDS budget of the rest-coordinate force", plan->max_u);
        return TLFEA_ECAPACITY;
    }
    return launch_force_rest(plan, x, y, z, X0, Y0, Z0, *quad, *mat, v, f, accumulate, nullptr, as_stream(stream));
}

extern "C" int tlfea_fixed_dof_mask(int n_nodes, int n_fixed, const int* fixed_nodes, unsigned char* mask,
                                    void* stream) {
    TLFEA_REQUIRE(n_nodes > 0 && mask && n_fixed >= 0 && (n_fixed == 0 || fixed_nodes), "bad argument");
    hipStream_t s = as_stream(stream);
    TLFEA_HIP_TRY(hipMemsetAsync(mask, 0, 3 * (size_t)n_nodes, s));
    if (n_fixed > 0) {
        k_fixed_mask<<<grid_for(n_fixed, 256), 256, 0, s>>>(n_fixed, fixed_nodes, mask);
        TLFEA_LAUNCH_CHECK();
    }
    return TLFEA_OK;
}
