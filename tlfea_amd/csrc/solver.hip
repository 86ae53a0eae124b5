// Implicit-Euler velocity solvers: SyncedNesterov and SyncedAdamW (cooperative and
// no-cooperative semantics).
//
// One inner iteration is four launches: a per-node predict kernel (momentum / Adam
// update, scratch positions x = x_prev + dt v, f_int = 0), the fused element force
// kernel, a per-row gradient kernel on the sliced mass layout (g, the next iterate and
// block partials of the norms) and a one-block finalize kernel that combines the
// partials in block order and sets the device convergence flag.  Every kernel returns at
// once when the flag is set, so the host may queue iterations ahead and reads the flag
// only every `poll` iterations.  Norms are tree reductions (the reference sums them
// serially on one thread, SyncedNesterov.cu:226-271), and every per-DOF phase covers all
// DOFs at any mesh size (the reference's cooperative grid is clamped, :392-397).
#include <cmath>
#include <vector>

#include "internal.h"

namespace tlfea_hip {
constexpr int S_RED = 256;  // threads per block of the reduction kernels
}


struct tlfea_solver {
    int kind = 0;
    tlfea_body b{};
    tlfea_nesterov_params np{};
    tlfea_adamw_params ap{};
    bool params_set = false;
    double rel_tol = 0.0;
    int poll = 16;
    int restart = 0;  // Nesterov gradient restart (tlfea_solver_set_restart)
    double mf_rho = 0.0;  // > 0: matrix-free mass of this density (tlfea_solver_set_matrix_free_mass)
    int64_t n3 = 0;
    // state
    double *v_guess = nullptr, *v_prev = nullptr, *lam = nullptr;
    double *xp = nullptr, *yp = nullptr, *zp = nullptr;
    double *vk = nullptr, *vn = nullptr;  // Nesterov iterates
    double *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    double* dv = nullptr;  // [n_nodes][4]: v_guess - v_prev per node, padded to 32 bytes
    double *ux = nullptr, *uy = nullptr, *uz = nullptr;  // displacements x - X0 (rest form)
    // constraints per node, built from the fixed list h_fixed
    int *cons_off = nullptr, *cons_idx = nullptr;
    unsigned char* node_fixed = nullptr;
    std::vector<int> h_fixed;
    // reductions
    double* part = nullptr;  // [5][max_blocks]
    int max_blocks = 0;
    double* scal = nullptr;  // device scalars, see k_finalize
    int* flags = nullptr;    // 0 inner_flag, 1 outer_flag, 2 inner count, 3 outer count
    // stats
    int last_inner = 0, last_outer = 0, last_restarts = 0;
    double last_norm_g = 0.0, last_rel = 0.0;
};

namespace tlfea_hip {

using Body = tlfea_body;
// one node's v_guess - v_prev, padded so that a lane's gather is two 16-byte loads
typedef double dv4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sums NV values over the block; results valid in thread 0
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV]) {
    __shared__ double red[NV][S_RED / 64];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wsum(v[k]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) red[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
            for (int j = 0; j < (int)(blockDim.x >> 6); j++) s += red[k][j];
            v[k] = s;
        }
    __syncthreads();
}

// ------------------------------------------------------------------------------------
// Nesterov
// ------------------------------------------------------------------------------------
// outer start: v_k = v_km1 = v_guess (SyncedNesterov.cu:129-134)
__global__ void k_nes_outer_init(int64_t n3, const double* __restrict__ vg, double* __restrict__ vk,
                                 double* __restrict__ vn, const int* __restrict__ flags, double* __restrict__ scal) {
    if (flags[1]) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) {
        vk[i] = vg[i];
        vn[i] = vg[i];
    }
    if (i == 0) {
        scal[0] = 0.0;  // prev_norm_g
        scal[1] = 0.0;  // norm_g
        scal[4] = -1.0; // norm_g_free at first inner iteration
        scal[11] = 1.0;                       // t_k of the device t-sequence (restart mode)
        scal[12] = 0.5 * (1.0 + sqrt(5.0));   // t_(k+1)
    }
}

// y = v_k + beta (v_k - v_km1); v_guess = y; x = x_prev + dt y; f_int = 0 (f_int null: the
// force launch overwrites it); rotate v_km1 <- v_k (SyncedNesterov.cu:152-175).  vk holds
// v_km1 on entry and v_k on exit; vn holds v_k.
// dv = y - v_prev, the same difference the gradient's mass term forms per entry
// (SyncedNesterov.cu:54-55), formed once per node for the gradient kernel's gathers.
// With tseq (restart mode) beta = (t_k - 1) / t_(k+1) comes from the device t-sequence
// (scal[11], scal[12]) instead of the argument.
// Displacements (u non-null, rest positions X0 known): u = (x_prev - X0) + dt y as well, for
// the force kernel: x_prev - X0 is exact while x_prev stays within a factor 2 of X0
// (Sterbenz), so u carries the rounding of its own size, not that of the coordinates.
struct DispOut {
    double *ux, *uy, *uz;
    const double *X0, *Y0, *Z0;
};
__device__ __forceinline__ void write_positions(int64_t i, int d, double dt, double yv, const double* __restrict__ xp,
                                                const double* __restrict__ yp, const double* __restrict__ zp,
                                                double* __restrict__ x, double* __restrict__ y,
                                                double* __restrict__ z, const DispOut& u) {
    double* const xo = d == 0 ? x : (d == 1 ? y : z);
    const double* const xq = d == 0 ? xp : (d == 1 ? yp : zp);
    xo[i] = xq[i] + dt * yv;
    if (u.ux) {
        double* const uo = d == 0 ? u.ux : (d == 1 ? u.uy : u.uz);
        const double* const Xq = d == 0 ? u.X0 : (d == 1 ? u.Y0 : u.Z0);
        uo[i] = (xq[i] - Xq[i]) + dt * yv;
    }
}

__global__ void k_nes_predict(int n_nodes, double beta, const double* __restrict__ tseq, double dt,
                              double* __restrict__ vk, const double* __restrict__ vn, double* __restrict__ vg,
                              const double* __restrict__ vp, double* __restrict__ dv, const double* __restrict__ xp,
                              const double* __restrict__ yp, const double* __restrict__ zp, double* __restrict__ x,
                              double* __restrict__ y, double* __restrict__ z, DispOut u, double* __restrict__ f_int,
                              const int* __restrict__ flags) {
    if (flags[0] | flags[1]) return;
    // one thread per DOF: the per-DOF streams are read and written coalesced
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * (int64_t)n_nodes) return;
    if (tseq) beta = (tseq[11] - 1.0) / tseq[12];
    const int64_t i = k / 3;
    const int d = (int)(k - 3 * i);
    const double vkk = vn[k], km1 = vk[k];
    const double yv = vkk + beta * (vkk - km1);
    vg[k] = yv;
    vk[k] = vkk;
    if (f_int) f_int[k] = 0.0;
    dv[4 * i + d] = yv - vp[k];  // component d of the node's padded difference
    write_positions(i, d, dt, yv, xp, yp, zp, x, y, z, u);
}

// Per node row on the sliced mass layout:
//   g = M (v_guess - v_prev)/dt + f_int - f_ext + dt sum_{c at this DOF} (lambda_c + rho c_c)
// (SyncedNesterov.cu:32-95), with c = x - x_ref at fixed nodes (FEAT10DataFunc.cuh:470-481).
// mode 0 (Nesterov): v_next = y - alpha g -> vn; partials |g|^2, |g_free|^2, |v_next|^2, |v_k|^2
//   and g.(v_next - v_k) (the restart test).
// mode 1 (AdamW): g -> gbuf; partials |g|^2, |g_free|^2, |v_guess|^2.
// The mass term gathers dv = v_guess - v_prev per column (written by the predict kernels),
// so each entry adds mass_ij * dv_j * inv_dt exactly as the per-entry difference would.
template <int MODE>
__global__ __launch_bounds__(S_RED) void k_grad(Body b, int64_t n_slices, const int* __restrict__ sell_row,
                                                const int64_t* __restrict__ slice_base,
                                                const int* __restrict__ slice_len, const int* __restrict__ sell_col,
                                                const double* __restrict__ vg, const dv4* __restrict__ dv,
                                                const double* __restrict__ lam, const int* __restrict__ cons_off,
                                                const int* __restrict__ cons_idx,
                                                const unsigned char* __restrict__ node_fixed, double dt, double rho,
                                                double alpha, const double* __restrict__ vk, double* __restrict__ out,
                                                double* __restrict__ part, int nblk, const int* __restrict__ flags) {
    if (flags[0] | flags[1]) return;
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t t = (int64_t)blk * blockDim.x + threadIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < n_slices * TILE) {
        const int row = sell_row[t];
        const int64_t sl = t >> 6;
        const int lane = (int)(t & (TILE - 1));
        const int64_t base = slice_base[sl];
        const int L = slice_len[sl];
        const double inv_dt = 1.0 / dt;
        double r0 = 0.0, r1 = 0.0, r2 = 0.0;
        const int* cp = sell_col + base + lane;
        const double* mp = b.m_sell + base + lane;
        int k = 0;
        // four entries' loads in flight at once; the sums still run in entry order
        for (; k + 4 <= L; k += 4) {
            int col[4];
            double mv[4];
            dv4 q[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                col[u] = cp[(int64_t)(k + u) * TILE];
                mv[u] = mp[(int64_t)(k + u) * TILE];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) q[u] = dv[col[u]];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                r0 += mv[u] * q[u].x * inv_dt;
                r1 += mv[u] * q[u].y * inv_dt;
                r2 += mv[u] * q[u].z * inv_dt;
            }
        }
        for (; k < L; k++) {
            const dv4 q = dv[cp[(int64_t)k * TILE]];
            const double mv = mp[(int64_t)k * TILE];
            r0 += mv * q.x * inv_dt;
            r1 += mv * q.y * inv_dt;
            r2 += mv * q.z * inv_dt;
        }
        if (row >= 0) {
            double gr[3] = {r0, r1, r2};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                gr[d] -= (-b.f_int[3 * row + d]);
                gr[d] -= b.f_ext[3 * row + d];
            }
            for (int k = cons_off[row]; k < cons_off[row + 1]; k++) {
                const int ci = cons_idx[k];
                const double xs[3] = {b.x[row], b.y[row], b.z[row]};
                const double xr[3] = {b.xref[row], b.yref[row], b.zref[row]};
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double c = xs[d] - xr[d];
                    gr[d] += dt * 1.0 * (lam[3 * ci + d] + rho * c);
                }
            }
            const bool fixed = node_fixed[row] != 0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                acc[0] += gr[d] * gr[d];
                if (!fixed) acc[1] += gr[d] * gr[d];
            }
            if (MODE == 0) {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double vnx = vg[3 * row + d] - alpha * gr[d];
                    out[3 * row + d] = vnx;
                    acc[2] += vnx * vnx;
                    const double k = vk[3 * row + d];
                    acc[3] += k * k;
                    acc[4] += gr[d] * (vnx - k);
                }
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    out[3 * row + d] = gr[d];
                    acc[2] += vg[3 * row + d] * vg[3 * row + d];
                }
            }
        }
    }
    block_sums<5>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 5; k++) part[k * nblk + blk] = acc[k];
}

// The same gradient with the mass term applied without the assembled matrix: the block
// partials of the rest-coordinate force kernel carry f_int + (rho/dt) M dv (the quadrature
// rule's mass of mass_matrix_qp_kernel, FEAT10Data.cu:200-271, per element), gathered per
// node here; one thread per node, the same five partial sums.
template <int MODE>
__global__ __launch_bounds__(S_RED) void k_grad_mf(Body b, const int* __restrict__ g_off,
                                                   const int* __restrict__ g_idx, const double* __restrict__ fpart,
                                                   const double* __restrict__ vg, const double* __restrict__ lam,
                                                   const int* __restrict__ cons_off, const int* __restrict__ cons_idx,
                                                   const unsigned char* __restrict__ node_fixed, double dt, double rho,
                                                   double alpha, const double* __restrict__ vk, double* __restrict__ out,
                                                   double* __restrict__ part, int nblk, const int* __restrict__ flags) {
    if (flags[0] | flags[1]) return;
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int row = (int)(blk * blockDim.x + threadIdx.x);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (row < b.n_nodes) {
        double gr[3] = {0.0, 0.0, 0.0};
        gather_slots3(g_idx, fpart, g_off[row], g_off[row + 1], gr[0], gr[1], gr[2]);
#pragma unroll
        for (int d = 0; d < 3; d++) gr[d] -= b.f_ext[3 * row + d];
        for (int k = cons_off[row]; k < cons_off[row + 1]; k++) {
            const int ci = cons_idx[k];
            const double xs[3] = {b.x[row], b.y[row], b.z[row]};
            const double xr[3] = {b.xref[row], b.yref[row], b.zref[row]};
#pragma unroll
            for (int d = 0; d < 3; d++) gr[d] += dt * 1.0 * (lam[3 * ci + d] + rho * (xs[d] - xr[d]));
        }
        const bool fixed = node_fixed[row] != 0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            acc[0] += gr[d] * gr[d];
            if (!fixed) acc[1] += gr[d] * gr[d];
        }
        if (MODE == 0) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double vnx = vg[3 * row + d] - alpha * gr[d];
                out[3 * row + d] = vnx;
                acc[2] += vnx * vnx;
                const double k = vk[3 * row + d];
                acc[3] += k * k;
                acc[4] += gr[d] * (vnx - k);
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                out[3 * row + d] = gr[d];
                acc[2] += vg[3 * row + d] * vg[3 * row + d];
            }
        }
    }
    block_sums<5>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 5; k++) part[k * nblk + blk] = acc[k];
}

// Combines the partials; sets flags.
//   scal: 0 prev_norm_g, 1 norm_g, 2 norm_v_next, 3 norm_v_k (or |v|), 4 gfree0, 5 gfree,
//         6 norm_g0 (AdamW), 7 rel
// kind 0 Nesterov: stop if inner>0 and (| |g|-|g_prev| | < tol or | |v_next|-|v_k| | < tol)
//   (SyncedNesterov.cu:226-271); prev_norm_g = norm_g.  With restart, advance the device
//   t-sequence (scal[11] = t_(k+1), scal[12] = t_(k+2)), or restart it at t_(k+1) = 1 when
//   g.(v_next - v_k) > 0 (the gradient test of O'Donoghue & Candes, Found. Comput. Math. 15
//   (2015) 715-732): the next step then carries no momentum.
// kind 1/2 AdamW (on check iterations only): stop if |g| <= tol (1+|v|) or
//   (rtol > 0 and |g| <= rtol |g0|) (SyncedAdamW.cu:238-272).
__global__ __launch_bounds__(S_RED) void k_finalize(int kind, int inner, int check, int nblk,
                                                    const double* __restrict__ part, double* __restrict__ scal,
                                                    int* __restrict__ flags, double inner_tol, double inner_rtol,
                                                    double rel_tol, int restart) {
    if (flags[0] | flags[1]) return;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nblk; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < 5; k++) s[k] += part[k * nblk + i];
    block_sums<5>(s);
    if (threadIdx.x != 0) return;
    const double ng = sqrt(s[0]), ngf = sqrt(s[1]);
    flags[2] += 1;
    if (scal[4] < 0.0) scal[4] = ngf;
    scal[5] = ngf;
    scal[7] = scal[4] > 0.0 ? ngf / scal[4] : 0.0;
    int stop = 0;
    if (kind == 0) {
        scal[1] = ng;
        const double nvn = sqrt(s[2]), nvk = sqrt(s[3]);
        if (inner > 0 && fabs(ng - scal[0]) < inner_tol) stop = 1;
        if (inner > 0 && fabs(nvn - nvk) < inner_tol) stop = 1;
        scal[0] = ng;
        scal[2] = nvn;
        scal[3] = nvk;
        if (restart) {
            const double t1 = s[4] > 0.0 ? 1.0 : scal[12];
            scal[11] = t1;
            scal[12] = 0.5 * (1.0 + sqrt(1.0 + 4.0 * t1 * t1));
            if (s[4] > 0.0) flags[6] += 1;  // restarts of this solve
        }
    } else if (check) {
        scal[1] = ng;
        if (scal[6] < 0.0) scal[6] = ng;
        const double nv = sqrt(s[2]);
        scal[3] = nv;
        const double tol_abs = inner_tol * (1.0 + nv);
        const double tol_rel = (inner_rtol > 0.0 && scal[6] > 0.0) ? inner_rtol * scal[6] : 0.0;
        if (ng <= tol_abs || (tol_rel > 0.0 && ng <= tol_rel)) stop = 1;
    } else {
        scal[1] = ng;
    }
    if (rel_tol > 0.0 && scal[7] <= rel_tol) stop = 1;
    if (stop) flags[0] = 1;
}

// ------------------------------------------------------------------------------------
// AdamW
// ------------------------------------------------------------------------------------
__global__ void k_adam_outer_init(int64_t n3, double* __restrict__ g, double* __restrict__ m1,
                                  double* __restrict__ m2, const int* __restrict__ flags, double* __restrict__ scal) {
    if (flags[1]) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) {
        g[i] = 0.0;
        m1[i] = 0.0;
        m2[i] = 0.0;
    }
    if (i == 0) {
        scal[1] = 0.0;
        scal[4] = -1.0;
        scal[6] = -1.0;  // norm_g0
    }
}

// m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; y = v_g - lr (m_hat/(sqrt(v_hat)+eps) + wd v_g)
// with m_hat = m * c1, v_hat = v * c2 (nocoop: c = 1/(1-beta^t) from the host) or
// m_hat = m / c1, v_hat = v / c2 (coop, SyncedAdamW.cu:146-167: c = 1 - pow(beta, t), the
// same for every DOF, so the host evaluates it once per iteration); then x = x_prev + dt y and
// f = 0 unless f_int is null (the force launch then overwrites it).
__global__ void k_adam_predict(int n_nodes, int nocoop, double lr, double c1, double c2, double beta1,
                               double beta2, double eps, double wd, double dt, const double* __restrict__ g,
                               double* __restrict__ m1, double* __restrict__ m2, double* __restrict__ vg,
                               const double* __restrict__ vp, double* __restrict__ dv, const double* __restrict__ xp,
                               const double* __restrict__ yp, const double* __restrict__ zp, double* __restrict__ x,
                               double* __restrict__ y, double* __restrict__ z, DispOut u, double* __restrict__ f_int,
                               const int* __restrict__ flags) {
    if (flags[0] | flags[1]) return;
    // one thread per DOF: the five per-DOF streams are read and written coalesced
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * (int64_t)n_nodes) return;
    const int64_t i = k / 3;
    const int d = (int)(k - 3 * i);
    const double gi = g[k], v0 = vg[k];
    double mt, vt, mh, vh;
    if (nocoop) {
        mt = beta1 * m1[k] + (1.0 - beta1) * gi;
        vt = beta2 * m2[k] + (1.0 - beta2) * gi * gi;
        mh = mt * c1;
        vh = vt * c2;
    } else {
        mt = beta1 * m1[k] + (1 - beta1) * gi;
        vt = beta2 * m2[k] + (1 - beta2) * gi * gi;
        mh = mt / c1;
        vh = vt / c2;
    }
    m1[k] = mt;
    m2[k] = vt;
    const double yv = v0 - lr * (mh / (sqrt(vh) + eps) + wd * v0);
    vg[k] = yv;
    if (f_int) f_int[k] = 0.0;
    dv[4 * i + d] = yv - vp[k];  // component d of the node's padded difference
    write_positions(i, d, dt, yv, xp, yp, zp, x, y, z, u);
}

// ------------------------------------------------------------------------------------
// outer-loop phases
// ------------------------------------------------------------------------------------
__global__ void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b, const int* __restrict__ flags,
                       int respect_outer) {
    if (respect_outer && flags[1]) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_save_prev(int n, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, double* __restrict__ xp, double* __restrict__ yp,
                            double* __restrict__ zp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    xp[i] = x[i];
    yp[i] = y[i];
    zp[i] = z[i];
}

// v_prev = v_guess; x = x_prev + dt v_guess (SyncedNesterov.cu:294-311)
__global__ void k_outer_positions(int n_nodes, double dt, const double* __restrict__ vg, double* __restrict__ vp,
                                  const double* __restrict__ xp, const double* __restrict__ yp,
                                  const double* __restrict__ zp, double* __restrict__ x, double* __restrict__ y,
                                  double* __restrict__ z, const int* __restrict__ flags, int set_vprev) {
    if (flags[1]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    if (set_vprev)
#pragma unroll
        for (int d = 0; d < 3; d++) vp[3 * i + d] = vg[3 * i + d];
    x[i] = xp[i] + vg[3 * i] * dt;
    y[i] = yp[i] + vg[3 * i + 1] * dt;
    z[i] = zp[i] + vg[3 * i + 2] * dt;
}

// c = x - x_ref at the fixed nodes; lambda += rho dt c; partial |c|^2 (one block)
// (SyncedNesterov.cu:315-342).
__global__ __launch_bounds__(S_RED) void k_dual(Body b, double rho, double dt, double* __restrict__ lam,
                                                double* __restrict__ scal, int* __restrict__ flags, double outer_tol,
                                                int nocoop) {
    if (flags[1]) return;
    double s[1] = {0.0};
    for (int k = threadIdx.x; k < b.n_fixed; k += blockDim.x) {
        const int n = b.fixed_nodes[k];
        const double c[3] = {b.x[n] - b.xref[n], b.y[n] - b.yref[n], b.z[n] - b.zref[n]};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            b.constraint[3 * k + d] = c[d];
            lam[3 * k + d] += rho * dt * c[d];
            s[0] += c[d] * c[d];
        }
    }
    block_sums<1>(s);
    if (threadIdx.x == 0) {
        const double nc = sqrt(s[0]);
        scal[8] = nc;
        flags[3] += 1;
        // nocoop: the outer stop also needs a converged inner loop (SyncedAdamWNocoop.cu:421-431)
        if (nc < outer_tol && (!nocoop || flags[0])) flags[1] = 1;
        flags[0] = 0;
    }
}

__global__ void k_constraint_eval(int n_fixed, const int* __restrict__ fixed, const double* __restrict__ x,
                                  const double* __restrict__ y, const double* __restrict__ z,
                                  const double* __restrict__ xr, const double* __restrict__ yr,
                                  const double* __restrict__ zr, double* __restrict__ c) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_fixed) return;
    const int n = fixed[k];
    c[3 * k] = x[n] - xr[n];
    c[3 * k + 1] = y[n] - yr[n];
    c[3 * k + 2] = z[n] - zr[n];
}

__global__ void k_reset_inner(int* __restrict__ flags) {
    if (threadIdx.x == 0 && blockIdx.x == 0) flags[0] = 0;
}

}  // namespace tlfea_hip

using namespace tlfea_hip;

// ======================================================================================
// C-ABI
// ======================================================================================
namespace {
template <typename T>
int dalloc(T** p, int64_t n) {
    *p = nullptr;
    TLFEA_HIP_TRY(hipMalloc((void**)p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    TLFEA_HIP_TRY(hipMemset(*p, 0, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    return TLFEA_OK;
}

void free_solver(tlfea_solver* s) {
    if (!s) return;
    for (void* p : {(void*)s->v_guess, (void*)s->v_prev, (void*)s->lam, (void*)s->xp, (void*)s->yp, (void*)s->zp,
                    (void*)s->vk, (void*)s->vn, (void*)s->g, (void*)s->m1, (void*)s->m2, (void*)s->dv, (void*)s->ux,
                    (void*)s->uy, (void*)s->uz, (void*)s->cons_off,
                    (void*)s->cons_idx, (void*)s->node_fixed, (void*)s->part, (void*)s->scal, (void*)s->flags})
        (void)hipFree(p);
    delete s;
}

// Rest positions in the body descriptor select the rest-coordinate force kernel: it needs
// the force plan, the standard rule and LDS room for the plan's largest block.
int check_rest_form(const tlfea_body* b) {
    if (!b->X0 && !b->Y0 && !b->Z0) return TLFEA_OK;
    TLFEA_REQUIRE(b->X0 && b->Y0 && b->Z0 && b->force_plan, "rest positions need all three arrays and a force plan");
    TLFEA_REQUIRE(is_standard_rule(b->nn, b->quad), "rest positions need the standard quadrature rule");
    const bool damp = b->mat.eta_damp != 0.0 || b->mat.lambda_damp != 0.0;
    if ((size_t)b->force_plan->max_u * (damp ? 12 : 9) * sizeof(double) > 160 * 1024) {
        set_error("a block touches %d nodes: beyond the LDS budget of the rest-coordinate force",
                  b->force_plan->max_u);
        return TLFEA_ECAPACITY;
    }
    return TLFEA_OK;
}

// Node -> constraint list (J^T in CSR over nodes, FEAT10Data.cu:420-473) and the per-node
// fixed flag, from the body's fixed-node list.  Called at create and again whenever the
// list changes (the drop-in's UpdateNodalFixed, FEAT10Data.cu:674-751): the gradient's
// J^T(lambda + rho c) term and the free-DOF norm then follow the new nodes.
int build_constraint_tables(tlfea_solver* s, const tlfea_body& body) {
    std::vector<int> fixed(body.n_fixed);
    if (body.n_fixed > 0 &&
        hipMemcpy(fixed.data(), body.fixed_nodes, sizeof(int) * body.n_fixed, hipMemcpyDeviceToHost) != hipSuccess) {
        tlfea_hip::set_error("could not read the fixed-node list");
        return TLFEA_EHIP;
    }
    std::vector<int> cnt(body.n_nodes + 1, 0), idx(body.n_fixed);
    std::vector<unsigned char> nf(body.n_nodes, 0);
    for (int k = 0; k < body.n_fixed; k++) {
        if (fixed[k] < 0 || fixed[k] >= body.n_nodes) {
            tlfea_hip::set_error("fixed node %d out of range", fixed[k]);
            return TLFEA_EINVAL;
        }
        cnt[fixed[k] + 1]++;
        nf[fixed[k]] = 1;
    }
    for (int i = 0; i < body.n_nodes; i++) cnt[i + 1] += cnt[i];
    std::vector<int> cur(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < body.n_fixed; k++) idx[cur[fixed[k]]++] = k;
    (void)hipFree(s->cons_idx);
    s->cons_idx = nullptr;
    if (int rc = dalloc(&s->cons_idx, body.n_fixed)) return rc;
    if (hipMemcpy(s->cons_off, cnt.data(), sizeof(int) * (body.n_nodes + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (body.n_fixed > 0 &&
         hipMemcpy(s->cons_idx, idx.data(), sizeof(int) * body.n_fixed, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(s->node_fixed, nf.data(), body.n_nodes, hipMemcpyHostToDevice) != hipSuccess) {
        tlfea_hip::set_error("copy of the constraint tables failed");
        return TLFEA_EHIP;
    }
    s->h_fixed.swap(fixed);
    return TLFEA_OK;
}
}  // namespace

extern "C" int tlfea_solver_create(int kind, const tlfea_body* body, tlfea_solver** out) {
    TLFEA_REQUIRE(out && body, "null argument");
    *out = nullptr;
    TLFEA_REQUIRE(kind >= 0 && kind <= 2, "unknown solver kind %d", kind);
    TLFEA_REQUIRE(body->nn == 10 || body->nn == 4, "unsupported element");
    TLFEA_REQUIRE_MODEL(body->mat);
    TLFEA_REQUIRE(body->pattern && body->m_sell && body->f_ext && body->f_int && body->x && body->y && body->z,
                  "body is missing device arrays");
    TLFEA_REQUIRE(body->pattern->n_nodes == body->n_nodes, "pattern does not match the body");
    TLFEA_REQUIRE(!body->force_plan || (body->force_plan->n_elem == body->n_elem && body->force_plan->nn == body->nn),
                  "force plan does not match the body");
    TLFEA_REQUIRE(body->n_fixed == 0 || (body->fixed_nodes && body->constraint && body->xref && body->yref && body->zref),
                  "constraints declared without arrays");
    if (int rc = check_rest_form(body)) return rc;
    auto* s = new tlfea_solver();
    s->kind = kind;
    s->b = *body;
    s->n3 = 3 * (int64_t)body->n_nodes;
    const int64_t n3 = s->n3;
    const int nblk = (int)((body->pattern->n_slices * TILE + S_RED - 1) / S_RED);
    s->max_blocks = nblk;
    int rc = TLFEA_OK;
    if ((rc = dalloc(&s->v_guess, n3)) || (rc = dalloc(&s->v_prev, n3)) || (rc = dalloc(&s->lam, 3 * body->n_fixed)) ||
        (rc = dalloc(&s->xp, body->n_nodes)) || (rc = dalloc(&s->yp, body->n_nodes)) ||
        (rc = dalloc(&s->zp, body->n_nodes)) || (rc = dalloc(&s->vk, n3)) || (rc = dalloc(&s->vn, n3)) ||
        (rc = dalloc(&s->g, n3)) || (rc = dalloc(&s->m1, n3)) || (rc = dalloc(&s->m2, n3)) || (rc = dalloc(&s->dv, 4 * (int64_t)body->n_nodes)) ||
        (rc = dalloc(&s->part, 5 * (int64_t)nblk)) || (rc = dalloc(&s->scal, 16)) || (rc = dalloc(&s->flags, 8)) ||
        (rc = dalloc(&s->ux, body->n_nodes)) || (rc = dalloc(&s->uy, body->n_nodes)) ||
        (rc = dalloc(&s->uz, body->n_nodes)) ||
        (rc = dalloc(&s->node_fixed, body->n_nodes)) || (rc = dalloc(&s->cons_off, body->n_nodes + 1))) {
        free_solver(s);
        return rc;
    }
    if ((rc = build_constraint_tables(s, *body))) {
        free_solver(s);
        return rc;
    }
    *out = s;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_destroy(tlfea_solver* s) {
    free_solver(s);
    return TLFEA_OK;
}

extern "C" int tlfea_solver_update_body(tlfea_solver* s, const tlfea_body* body) {
    TLFEA_REQUIRE(s && body, "null argument");
    TLFEA_REQUIRE(body->nn == s->b.nn && body->n_elem == s->b.n_elem && body->n_nodes == s->b.n_nodes,
                  "body sizes differ from the solver's");
    TLFEA_REQUIRE(body->pattern && body->m_sell && body->f_ext && body->f_int, "body is missing device arrays");
    TLFEA_REQUIRE(body->n_fixed == 0 || (body->fixed_nodes && body->constraint && body->xref && body->yref && body->zref),
                  "constraints declared without arrays");
    TLFEA_REQUIRE_MODEL(body->mat);
    if (int rc = check_rest_form(body)) return rc;
    // a changed fixed set (UpdateNodalFixed) rebuilds the node -> constraint tables.  With
    // the same count the multipliers are kept, indexed by constraint number as the
    // reference's solver keeps d_lambda_guess; a new count starts them at zero.
    std::vector<int> fixed(body->n_fixed);
    if (body->n_fixed > 0)
        TLFEA_HIP_TRY(hipMemcpy(fixed.data(), body->fixed_nodes, sizeof(int) * body->n_fixed, hipMemcpyDeviceToHost));
    if (body->n_fixed != s->b.n_fixed) {
        (void)hipFree(s->lam);
        s->lam = nullptr;
        if (int rc = dalloc(&s->lam, 3 * (int64_t)body->n_fixed)) return rc;
    }
    if (fixed != s->h_fixed)
        if (int rc = build_constraint_tables(s, *body)) return rc;
    s->b = *body;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_set_parameters(tlfea_solver* s, const void* params) {
    TLFEA_REQUIRE(s && params, "null argument");
    if (s->kind == TLFEA_SOLVER_NESTEROV)
        s->np = *static_cast<const tlfea_nesterov_params*>(params);
    else
        s->ap = *static_cast<const tlfea_adamw_params*>(params);
    s->params_set = true;
    TLFEA_HIP_TRY(hipMemset(s->v_guess, 0, sizeof(double) * s->n3));
    TLFEA_HIP_TRY(hipMemset(s->v_prev, 0, sizeof(double) * s->n3));
    TLFEA_HIP_TRY(hipMemset(s->lam, 0, sizeof(double) * (size_t)(3 * s->b.n_fixed > 0 ? 3 * s->b.n_fixed : 1)));
    return TLFEA_OK;
}

extern "C" int tlfea_solver_set_rel_tol(tlfea_solver* s, double rel_tol) {
    TLFEA_REQUIRE(s && rel_tol >= 0.0, "bad argument");
    s->rel_tol = rel_tol;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_set_restart(tlfea_solver* s, int on) {
    TLFEA_REQUIRE(s, "null solver");
    TLFEA_REQUIRE(on == 0 || s->kind == TLFEA_SOLVER_NESTEROV, "restart applies to the Nesterov solver only");
    s->restart = on != 0;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_set_matrix_free_mass(tlfea_solver* s, double rho) {
    TLFEA_REQUIRE(s != nullptr, "null solver");
    TLFEA_REQUIRE(rho >= 0.0, "density must be non-negative");
    if (rho > 0.0) {
        const tlfea_body& b = s->b;
        TLFEA_REQUIRE(b.force_plan && b.X0, "the matrix-free mass rides on the rest-coordinate force: a body "
                                            "with a force plan and rest positions is needed");
        TLFEA_REQUIRE(b.mat.model == 0 && b.mat.eta_damp == 0.0 && b.mat.lambda_damp == 0.0,
                      "the matrix-free mass covers SVK bodies without damping");
        TLFEA_REQUIRE(is_standard_rule(b.nn, b.quad), "the matrix-free mass needs the standard quadrature rule");
        if ((size_t)b.force_plan->max_u * 12 * sizeof(double) > 160 * 1024) {
            set_error("a block touches %d nodes: beyond the LDS budget of the force with a mass term",
                      b.force_plan->max_u);
            return TLFEA_ECAPACITY;
        }
    }
    s->mf_rho = rho;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_restarts(const tlfea_solver* s, int* restarts) {
    TLFEA_REQUIRE(s && restarts, "null argument");
    *restarts = s->last_restarts;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_set_poll_interval(tlfea_solver* s, int every) {
    TLFEA_REQUIRE(s && every >= 1, "bad argument");
    s->poll = every;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_state(tlfea_solver* s, double** v_guess, double** v_prev, double** lambda_guess) {
    TLFEA_REQUIRE(s, "null solver");
    if (v_guess) *v_guess = s->v_guess;
    if (v_prev) *v_prev = s->v_prev;
    if (lambda_guess) *lambda_guess = s->lam;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_stats(const tlfea_solver* s, int* inner_iters, int* outer_iters, double* norm_g,
                                  double* rel_g) {
    TLFEA_REQUIRE(s, "null solver");
    if (inner_iters) *inner_iters = s->last_inner;
    if (outer_iters) *outer_iters = s->last_outer;
    if (norm_g) *norm_g = s->last_norm_g;
    if (rel_g) *rel_g = s->last_rel;
    return TLFEA_OK;
}

extern "C" int tlfea_solver_solve(tlfea_solver* S, void* stream) {
    TLFEA_REQUIRE(S, "null solver");
    TLFEA_REQUIRE(S->params_set, "SetParameters must be called before Solve");
    hipStream_t st = as_stream(stream);
    const tlfea_body& b = S->b;
    const tlfea_pattern* P = b.pattern;
    const int N = b.n_nodes;
    const int64_t n3 = S->n3;
    const bool nes = S->kind == TLFEA_SOLVER_NESTEROV;
    const bool nocoop = S->kind == TLFEA_SOLVER_ADAMW_NOCOOP;
    const double dt = nes ? S->np.time_step : S->ap.time_step;
    const double rho = nes ? S->np.rho : S->ap.rho;
    const int max_inner = nes ? S->np.max_inner : S->ap.max_inner;
    int max_outer = nes ? S->np.max_outer : S->ap.max_outer;
    if (nocoop && b.n_fixed == 0) max_outer = 1;  // SyncedAdamWNocoop.cu:298-301
    const double inner_tol = nes ? S->np.inner_tol : S->ap.inner_tol;
    const double outer_tol = nes ? S->np.outer_tol : S->ap.outer_tol;
    int check_int = nes ? 1 : S->ap.convergence_check_interval;
    if (check_int <= 0) {
        TLFEA_REQUIRE(nocoop, "convergence_check_interval must be positive");
        check_int = 1;  // SyncedAdamWNocoop.cu:294-296
    }
    TLFEA_REQUIRE(dt != 0.0, "time step must be nonzero");
    const unsigned gN = grid_for(N, 256), g3 = grid_for(n3, 256);
    const unsigned gS = (unsigned)S->max_blocks;
    const bool damp = b.mat.eta_damp != 0.0 || b.mat.lambda_damp != 0.0;
    // the planned force launches overwrite f_int (accumulate = 0); the atomic one adds to it
    double* const fzero = b.force_plan ? nullptr : b.f_int;
    // the rest-coordinate force takes the displacements the predict kernels form
    const bool rest_form = b.force_plan && b.X0;
    const DispOut uo = rest_form ? DispOut{S->ux, S->uy, S->uz, b.X0, b.Y0, b.Z0} : DispOut{};
    const double* const tseq = (nes && S->restart) ? S->scal : nullptr;
    // matrix-free mass: set up for a rest-coordinate body (tlfea_solver_set_matrix_free_mass)
    const bool mf = S->mf_rho > 0.0 && rest_form;
    const unsigned gM = grid_for(N, S_RED);

    TLFEA_HIP_TRY(hipMemsetAsync(S->flags, 0, sizeof(int) * 8, st));
    k_save_prev<<<gN, 256, 0, st>>>(N, b.x, b.y, b.z, S->xp, S->yp, S->zp);
    TLFEA_LAUNCH_CHECK();
    int hflags[8] = {0};
    for (int outer = 0; outer < max_outer; outer++) {
        if (nes)
            k_nes_outer_init<<<g3, 256, 0, st>>>(n3, S->v_guess, S->vk, S->vn, S->flags, S->scal);
        else
            k_adam_outer_init<<<g3, 256, 0, st>>>(n3, S->g, S->m1, S->m2, S->flags, S->scal);
        TLFEA_LAUNCH_CHECK();
        double t = 1.0, lr = nes ? 0.0 : S->ap.lr;
        bool inner_done = false;
        for (int inner = 0; inner < max_inner && !inner_done; inner++) {
            if (nes) {
                const double t_next = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
                const double beta = (t - 1.0) / t_next;
                t = t_next;
                k_nes_predict<<<g3, 256, 0, st>>>(N, beta, tseq, dt, S->vk, S->vn, S->v_guess, S->v_prev, S->dv,
                                                  S->xp, S->yp, S->zp, b.x, b.y, b.z, uo, fzero, S->flags);
            } else {
                double lr_k, c1 = 0.0, c2 = 0.0;
                if (nocoop) {
                    lr_k = S->ap.lr * std::pow(S->ap.lr_decay, inner + 1);
                    const double tc = (double)(inner + 2);
                    c1 = 1.0 / (1.0 - std::pow(S->ap.beta1, tc));
                    c2 = 1.0 / (1.0 - std::pow(S->ap.beta2, tc));
                    t = tc;
                } else {
                    lr = lr * S->ap.lr_decay;
                    t += 1;
                    lr_k = lr;
                    c1 = 1 - std::pow(S->ap.beta1, t);
                    c2 = 1 - std::pow(S->ap.beta2, t);
                }
                k_adam_predict<<<g3, 256, 0, st>>>(N, nocoop ? 1 : 0, lr_k, c1, c2, S->ap.beta1, S->ap.beta2,
                                                   S->ap.eps, S->ap.weight_decay, dt, S->g, S->m1, S->m2, S->v_guess,
                                                   S->v_prev, S->dv, S->xp, S->yp, S->zp, b.x, b.y, b.z, uo,
                                                   fzero, S->flags);
            }
            TLFEA_LAUNCH_CHECK();
            // element forces at the scratch positions, damping from v_guess; skipped once the
            // inner loop has converged (f_int then keeps the last evaluation)
            if (mf) {
                ForceMass fm;
                fm.dv = S->dv;
                fm.dv_stride = 4;
                fm.mc = S->mf_rho / dt;
                fm.flags = S->flags;
                if (int rc = launch_force_rest_mass_partials(b.force_plan, S->ux, S->uy, S->uz, b.X0, b.Y0, b.Z0,
                                                             b.quad, b.mat, nullptr, nullptr, nullptr, fm, st))
                    return rc;
            } else if (rest_form) {
                if (int rc = launch_force_rest(b.force_plan, S->ux, S->uy, S->uz, b.X0, b.Y0, b.Z0, b.quad, b.mat,
                                               damp ? S->v_guess : nullptr, b.f_int, 0, S->flags, st, nullptr,
                                               nullptr, nullptr, true))
                    return rc;
            } else if (b.force_plan) {
                if (int rc = launch_force_planned(b.force_plan, b.x, b.y, b.z, b.grad_t, b.detJ_t, b.quad, b.mat,
                                                  damp ? S->v_guess : nullptr, b.f_int, 0, S->flags, st))
                    return rc;
            } else if (int rc = launch_force(b.nn, b.n_elem, b.conn, b.x, b.y, b.z, b.grad_t, b.detJ_t, b.quad,
                                             b.mat, damp ? S->v_guess : nullptr, b.f_int, S->flags, st)) {
                return rc;
            }
            const int check = nes ? 1 : ((inner % check_int) == 0);
            const tlfea_force_plan* fp = b.force_plan;
            if (mf && nes) {
                k_grad_mf<0><<<gM, S_RED, 0, st>>>(b, fp->d_g_off, fp->d_g_idx, fp->d_part, S->v_guess, S->lam,
                                                   S->cons_off, S->cons_idx, S->node_fixed, dt, rho, S->np.alpha,
                                                   S->vk, S->vn, S->part, gM, S->flags);
            } else if (mf) {
                k_grad_mf<1><<<gM, S_RED, 0, st>>>(b, fp->d_g_off, fp->d_g_idx, fp->d_part, S->v_guess, S->lam,
                                                   S->cons_off, S->cons_idx, S->node_fixed, dt, rho, 0.0, nullptr,
                                                   S->g, S->part, gM, S->flags);
            } else if (nes) {
                k_grad<0><<<gS, S_RED, 0, st>>>(b, P->n_slices, P->d_sell_row, P->d_slice_base, P->d_slice_len,
                                                P->d_sell_col, S->v_guess, reinterpret_cast<const dv4*>(S->dv), S->lam, S->cons_off, S->cons_idx,
                                                S->node_fixed, dt, rho, S->np.alpha, S->vk, S->vn, S->part, gS,
                                                S->flags);
            } else {
                k_grad<1><<<gS, S_RED, 0, st>>>(b, P->n_slices, P->d_sell_row, P->d_slice_base, P->d_slice_len,
                                                P->d_sell_col, S->v_guess, reinterpret_cast<const dv4*>(S->dv), S->lam, S->cons_off, S->cons_idx,
                                                S->node_fixed, dt, rho, 0.0, nullptr, S->g, S->part, gS, S->flags);
            }
            k_finalize<<<1, S_RED, 0, st>>>(S->kind, inner, check, mf ? gM : gS, S->part, S->scal, S->flags, inner_tol,
                                            nes ? 0.0 : S->ap.inner_rtol, S->rel_tol, tseq ? 1 : 0);
            TLFEA_LAUNCH_CHECK();
            if ((inner + 1) % S->poll == 0) {
                TLFEA_HIP_TRY(hipMemcpyAsync(hflags, S->flags, sizeof(int) * 8, hipMemcpyDeviceToHost, st));
                TLFEA_HIP_TRY(hipStreamSynchronize(st));
                inner_done = hflags[0] != 0;
            }
        }
        // with the matrix-free mass the partials held f_int + mass term: f_int itself at the
        // last evaluated positions (flags + 4: two always-zero words, so this launch runs)
        if (mf) {
            if (int rc = launch_force_rest(b.force_plan, S->ux, S->uy, S->uz, b.X0, b.Y0, b.Z0, b.quad, b.mat, nullptr,
                                           b.f_int, 0, S->flags + 4, st, nullptr, nullptr, nullptr, true))
                return rc;
        }
        // Nesterov keeps the last iterate in vn: v_guess = v_next (SyncedNesterov.cu:283)
        if (nes) k_copy<<<g3, 256, 0, st>>>(n3, S->vn, S->v_guess, S->flags, 1);
        k_outer_positions<<<gN, 256, 0, st>>>(N, dt, S->v_guess, S->v_prev, S->xp, S->yp, S->zp, b.x, b.y, b.z,
                                              S->flags, 1);
        TLFEA_LAUNCH_CHECK();
        if (!nocoop || b.n_fixed > 0) {
            k_dual<<<1, S_RED, 0, st>>>(b, rho, dt, S->lam, S->scal, S->flags, outer_tol, nocoop ? 1 : 0);
        } else {
            k_reset_inner<<<1, 1, 0, st>>>(S->flags);
        }
        TLFEA_LAUNCH_CHECK();
        TLFEA_HIP_TRY(hipMemcpyAsync(hflags, S->flags, sizeof(int) * 8, hipMemcpyDeviceToHost, st));
        TLFEA_HIP_TRY(hipStreamSynchronize(st));
        if (hflags[1]) break;  // outer converged: the remaining outer iterations do nothing
    }
    // final x = x_prev + dt v_guess (SyncedNesterov.cu:350-357); flags + 4 points at two
    // always-zero words, so this launch runs whatever the outer flag says
    k_outer_positions<<<gN, 256, 0, st>>>(N, dt, S->v_guess, S->v_prev, S->xp, S->yp, S->zp, b.x, b.y, b.z,
                                          S->flags + 4, 0);
    TLFEA_LAUNCH_CHECK();
    double hs[16];
    TLFEA_HIP_TRY(hipMemcpyAsync(hflags, S->flags, sizeof(int) * 8, hipMemcpyDeviceToHost, st));
    TLFEA_HIP_TRY(hipMemcpyAsync(hs, S->scal, sizeof(double) * 16, hipMemcpyDeviceToHost, st));
    TLFEA_HIP_TRY(hipStreamSynchronize(st));
    S->last_inner = hflags[2];
    S->last_outer = hflags[3];
    S->last_restarts = hflags[6];
    S->last_norm_g = hs[1];
    S->last_rel = hs[7];
    return TLFEA_OK;
}

extern "C" int tlfea_constraint_eval(int n_fixed, const int* fixed_nodes, const double* x, const double* y,
                                     const double* z, const double* xref, const double* yref, const double* zref,
                                     double* c, void* stream) {
    TLFEA_REQUIRE(n_fixed >= 0, "negative constraint count");
    if (n_fixed == 0) return TLFEA_OK;
    TLFEA_REQUIRE(fixed_nodes && x && y && z && xref && yref && zref && c, "null argument");
    k_constraint_eval<<<grid_for(n_fixed, 256), 256, 0, as_stream(stream)>>>(n_fixed, fixed_nodes, x, y, z, xref, yref,
                                                                           zref, c);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}
