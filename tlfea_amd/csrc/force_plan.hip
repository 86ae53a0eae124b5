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
#include <vector>

#include <cstring>

#include "internal.h"

namespace tlfea_hip {


template <int NN, int NQ, bool DAMP>
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
    const int b = blockIdx.x;
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
            svk_stress(F, mat.lambda, mat.mu, P);
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

// f[3i+d] (+)= sum over node i's partials, in ascending partial order (deterministic).
__global__ void k_gather_partials(int n_nodes, const int* __restrict__ g_off, const int64_t* __restrict__ g_idx,
                                  const double* __restrict__ part, double* __restrict__ f, int accumulate,
                                  const int* __restrict__ flags) {
    if (flags && (flags[0] | flags[1])) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = g_off[i]; k < g_off[i + 1]; k++) {
        const double* p = part + 3 * g_idx[k];
        s0 += p[0];
        s1 += p[1];
        s2 += p[2];
    }
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
        if (p->nn == 10) {
            if (damp)
                k_force_planned<10, 5, true><<<nb, PLAN_BLOCK, lds, s>>>(p->n_elem, p->d_blk_off, p->d_unodes,
                                                                         p->d_slot, x, y, z, v, grad_t, detJ_t, quad,
                                                                         mat, p->d_part, p->max_u, flags);
            else
                k_force_planned<10, 5, false><<<nb, PLAN_BLOCK, lds, s>>>(p->n_elem, p->d_blk_off, p->d_unodes,
                                                                          p->d_slot, x, y, z, nullptr, grad_t, detJ_t,
                                                                          quad, mat, p->d_part, p->max_u, flags);
        } else {
            if (damp)
                k_force_planned<4, 1, true><<<nb, PLAN_BLOCK, lds, s>>>(p->n_elem, p->d_blk_off, p->d_unodes,
                                                                        p->d_slot, x, y, z, v, grad_t, detJ_t, quad,
                                                                        mat, p->d_part, p->max_u, flags);
            else
                k_force_planned<4, 1, false><<<nb, PLAN_BLOCK, lds, s>>>(p->n_elem, p->d_blk_off, p->d_unodes,
                                                                         p->d_slot, x, y, z, nullptr, grad_t, detJ_t,
                                                                         quad, mat, p->d_part, p->max_u, flags);
        }
        TLFEA_LAUNCH_CHECK();
    }
    k_gather_partials<<<grid_for(p->n_nodes, 256), 256, 0, s>>>(p->n_nodes, p->d_g_off, p->d_g_idx, p->d_part, f,
                                                                 accumulate, flags);
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
    std::vector<int64_t> blk_off(nb + 1, 0);
    std::vector<int> unodes;
    unodes.reserve((size_t)n_elem * 3);
    std::vector<unsigned short> slot((size_t)n_tiles * nn * TILE, 0);
    std::vector<int> buf;
    int max_u = 0;
    for (int b = 0; b < nb; b++) {
        const int e0 = b * PLAN_BLOCK, e1 = std::min(n_elem, e0 + PLAN_BLOCK);
        buf.clear();
        for (int e = e0; e < e1; e++)
            for (int a = 0; a < nn; a++) buf.push_back(h_conn[e + (size_t)a * n_elem]);
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        const int U = (int)buf.size();
        max_u = std::max(max_u, U);
        for (int e = e0; e < e1; e++)
            for (int a = 0; a < nn; a++) {
                const int node = h_conn[e + (size_t)a * n_elem];
                const int sidx = (int)(std::lower_bound(buf.begin(), buf.end(), node) - buf.begin());
                slot[((size_t)(e >> 6) * nn + a) * TILE + (e & (TILE - 1))] = (unsigned short)sidx;
            }
        unodes.insert(unodes.end(), buf.begin(), buf.end());
        blk_off[b + 1] = blk_off[b] + U;
    }
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
    std::vector<int64_t> g_idx(unodes.size());
    std::vector<int> cur(g_off.begin(), g_off.end() - 1);
    for (int64_t k = 0; k < (int64_t)unodes.size(); k++) g_idx[cur[unodes[k]]++] = k;
    int rc = TLFEA_OK;
    if ((rc = upload(&P->d_blk_off, blk_off)) || (rc = upload(&P->d_unodes, unodes)) ||
        (rc = upload(&P->d_slot, slot)) || (rc = upload(&P->d_g_off, g_off)) || (rc = upload(&P->d_g_idx, g_idx))) {
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
    TLFEA_REQUIRE(quad->n == (plan->nn == 10 ? 5 : 1), "quadrature does not match the element");
    return launch_force_planned(plan, x, y, z, grad_t, detJ_t, *quad, *mat, v, f, accumulate, nullptr,
                                as_stream(stream));
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
