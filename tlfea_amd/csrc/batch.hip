// Many independent meshes solved together (not in the reference).
//
// The instances are concatenated into one element/node numbering (instance i owns the
// contiguous ranges elem_off[i]..elem_off[i+1] and node_off[i]..node_off[i+1]), so one
// launch of each kernel covers the whole batch.  Element kernels use the force plan of
// the concatenated mesh (a block may straddle two instances: nodes are never shared
// between instances, so a node's partials still come from its own instance only) and look
// up the instance's material per element.  Vector kernels run over "chunks": node ranges
// of at most chunk_nodes nodes inside one instance, one workgroup each, so that every
// dot product is reduced per chunk and then, in a fixed order, per instance.
//
// The linear-static CG is the reference's SolveLinearSystemCG
// (src/solvers/LinearStaticSolver.cu:348-443) run for every instance at once, with its
// own alpha/beta and stopping test per instance; finished instances stop updating and
// their elements are skipped.  The stopping test is ||f - K u|| / ||f|| < tol on the free
// DOFs (the reference's sqrt(rr/rr0) when the initial guess is zero).  An optional FP32
// Chebyshev preconditioner applies the same operator in single precision.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "internal.h"

// coarse-space modes per node plane: the 6 rigid motions, and with 11 also five in-plane
// fields (the quadratic cross-section warping of bending about either axis, uniform lateral
// stretch and the two in-plane shears)
constexpr int CM_MAX = 11;

struct tlfea_batch {
    int nn = 0, n_inst = 0, n_elem = 0, n_nodes = 0;
    tlfea_quad quad{};
    tlfea_force_plan* plan = nullptr;
    std::vector<int> h_elem_off, h_node_off;
    int* d_conn = nullptr;
    double *d_X = nullptr, *d_Y = nullptr, *d_Z = nullptr;
    unsigned char* d_mask = nullptr;
    int* d_elem_inst = nullptr;
    double *d_jinv = nullptr, *d_detJ = nullptr;
    float *d_jinv32 = nullptr, *d_detJ32 = nullptr;
    int affine = 0;  // every element's M and det J equal at all quadrature points
    double *d_lam = nullptr, *d_mu = nullptr;
    float *d_lam32 = nullptr, *d_mu32 = nullptr;
    // chunk tables
    int chunk_nodes = 0, n_chunks = 0;
    int *d_chunk_inst = nullptr, *d_chunk_beg = nullptr, *d_chunk_end = nullptr, *d_inst_chunk_off = nullptr;
    // solver workspace
    double *d_r = nullptr, *d_p = nullptr, *d_q = nullptr, *d_z = nullptr;
    float *d_r32 = nullptr, *d_z32 = nullptr, *d_d32 = nullptr, *d_q32 = nullptr;
    double* d_part = nullptr;  // [n_chunks * 2]
    double* d_s = nullptr;     // [n_inst * NS] per-instance scalars
    int* d_i = nullptr;        // [n_inst * NI] per-instance flags
    int* d_all = nullptr;      // [4]: 0 active instances
    double* d_cheb = nullptr;  // [n_inst * 4]: lambda_max estimate, theta, delta, sigma1
    double* d_rold = nullptr;  // previous residual (flexible CG)
    float* d_dinv32 = nullptr;  // inverse diagonal of K (1 on fixed DOFs)
    int cheb_degree = 0;        // 0 = no preconditioner set up
    double* d_mc = nullptr;     // per-instance mass coefficient (rho * c) of the operator K + c M
    float* d_mc32 = nullptr;    // the same in FP32, for the preconditioner's operator
    bool mass_on = false;
    int mass_exact = 0;         // 0: quadrature-rule mass, 1: exactly integrated mass
    // current displacement from the rest positions (tlfea_batch_set_displacement); the
    // operator is K(X + u) there, K at rest while moved is false
    double *d_ux = nullptr, *d_uy = nullptr, *d_uz = nullptr;
    bool moved = false;
    // active-work lists (used once finished instances are compacted out of the launches)
    std::vector<int> h_inst_chunk_off;
    int* d_blk_list = nullptr;
    int* d_chunk_list = nullptr;
    int* d_inst_list = nullptr;
    bool compact = false;
    int n_blk_list = 0, n_chunk_list = 0, n_inst_list = 0;
    int64_t launched_elems = 0;
    double compact_ratio = 0.6;  // compact when active elements fall below this share
    int poll_default = 32;       // iterations per graph replay when a solve passes poll <= 0
    hipStream_t s_own = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0, graph_precond = -1;
    double graph_tol = -1.0;
    // the CG iterates on this buffer (the caller's u is copied in and out), so a captured
    // graph stays valid from one solve to the next while the lists, poll, precond, tol and
    // the set-up it bakes in are unchanged
    double* d_usol = nullptr;
    // two-level coarse correction (tlfea_batch_coarse_setup): 6 rigid modes per node plane
    // across each instance's long axis; coarse DOF 6 p + j of (global, padded) plane p;
    // blocks of W planes (NB = 6 W coarse DOFs) form a block-tridiagonal matrix
    bool coarse_on = false;
    int cw = 0, cnb = 0, n_planes = 0, n_cblocks = 0;
    int cm_req = 6;  // modes per plane requested (tlfea_batch_set_coarse_modes)
    int cm = 6;      // modes per plane in use: cm_req, or the 6 rigid ones when W > 2
    int* d_plane_axis = nullptr;   // [n_planes]: the plane's normal (its instance's long axis)
    int* d_node_plane = nullptr;   // [n_nodes]: global plane, -1 for fixed nodes
    double* d_node_r = nullptr;    // [3 n_nodes]: (X - centroid) / size of the node's plane
    int *d_plane_off = nullptr, *d_plane_nodes = nullptr;  // plane -> free nodes (CSR)
    int* d_plane_inst = nullptr;   // [n_planes]
    int* d_plane_local = nullptr;  // [n_planes]: plane index within its instance
    int* d_inst_cblk = nullptr;    // [n_inst + 1]: first coarse block of each instance
    double *d_cdiag = nullptr, *d_clow = nullptr, *d_cg = nullptr;  // [n_cblocks][NB][NB]
    double *d_crhs = nullptr, *d_cx = nullptr;                       // [cm n_planes]
    // instances whose coarse matrix is applied as a dense inverse (long block chains, where
    // the sweep's latency exceeds reading the inverse): offsets into d_cdense (-1 = sweep),
    // and their coarse rows (global coarse DOF, instance)
    int64_t* d_cdense_off = nullptr;
    double* d_cdense = nullptr;
    int *d_cd_row = nullptr, *d_cd_inst = nullptr;
    int n_cd_rows = 0, n_cd_inst = 0;
    std::vector<int64_t> h_cdense_off;  // per instance (-1 = sweep)
    int* d_cinst_pn = nullptr;          // padded plane count per instance
    std::vector<int> h_inst_cblk;
    bool coarse_geom = false;           // the geometric tables above are built
    bool restrict_wide = false;         // planes of 512 free nodes or more on average
    // batched SyncedNesterov (tlfea_batch_nesterov_solve): v_(k-1) (AoS [3 n_nodes]), the
    // instance of every node, chunk partials [n_chunks][4], per-instance scalars
    // [n_inst][NSN], mass coefficients rho_i / dt, restart counts
    double* d_nvk = nullptr;
    int* d_node_inst = nullptr;
    double *d_npart = nullptr, *d_ns = nullptr, *d_nmc = nullptr;
    int* d_nrs = nullptr;
    int n_npart = 0;
};

namespace tlfea_hip {

constexpr int BT = 256;  // threads per chunk workgroup
// per-instance scalars
// (S_G0/S_G1, S_A0/S_A1: gamma and alpha of the single-reduction CG by iteration parity)
enum { S_FF = 0, S_RZ, S_RR, S_ALPHA, S_BETA, S_PQ, S_REL, S_RZ_OLD, S_G0, S_G1, S_A0, S_A1, NS };
// Chebyshev parameters per instance
enum { C_LMAX = 0, C_THETA, C_DELTA, C_SIGMA, NC };
// per-instance flags
// (I_RESET, I_HALVE: the batched Nesterov loop's restart-from-v_n request and step halvings)
enum { I_DONE = 0, I_ITERS, I_NOTPD, I_ZERO_F, I_RESET, I_HALVE, NI };

__device__ __forceinline__ double chunk_sum(double v) {
    __shared__ double red[BT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BT / 64; w++) t += red[w];
    return t;
}

struct Chunks {
    const int* inst;
    const int* beg;
    const int* end;
    const int* list;  // launched chunk ids (null = all)
};
__device__ __forceinline__ int chunk_id(const Chunks& ch) { return ch.list ? ch.list[blockIdx.x] : blockIdx.x; }

// instance ids handled by the per-instance reduction kernels (null list = 0..n-1)
struct Insts {
    int n;
    const int* list;
};
__device__ __forceinline__ int inst_at(const Insts& in, int j) { return in.list ? in.list[j] : j; }

// The K chunk partials cpart[K c + k] summed over instance i's chunks by one wave, in a
// fixed order (lane l takes chunks l, l + 64, ..., then a butterfly); lane 0 holds the
// sums.  The per-instance reduction kernels run one wave per instance.
template <int K>
__device__ __forceinline__ void wave_inst_sums(const int* __restrict__ ico, const double* __restrict__ cpart, int i,
                                               double (&s)[K]) {
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int c = ico[i] + (int)threadIdx.x; c < ico[i + 1]; c += 64)
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += cpart[K * (int64_t)c + k];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_xor(s[k], o, 64);
}

// q = sum of partials per node (or p on masked DOFs); chunk partial of p.q
template <typename T>
__global__ __launch_bounds__(BT) void k_bgather(Chunks ch, const int* __restrict__ g_off,
                                                const int* __restrict__ g_idx, const T* __restrict__ part,
                                                const unsigned char* __restrict__ mask, const T* __restrict__ p,
                                                T* __restrict__ q, double* __restrict__ cpart,
                                                const int* __restrict__ iflags) {
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    if (iflags && iflags[inst * NI + I_DONE]) {
        if (threadIdx.x == 0 && cpart) cpart[c] = 0.0;
        return;
    }
    double dot = 0.0;
    for (int i = ch.beg[c] + threadIdx.x; i < ch.end[c]; i += BT) {
        T s0 = T(0), s1 = T(0), s2 = T(0);
        gather_slots3(g_idx, part, g_off[i], g_off[i + 1], s0, s1, s2);
        const T sv[3] = {s0, s1, s2};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T v = (mask && mask[3 * i + d]) ? p[3 * i + d] : sv[d];
            q[3 * i + d] = v;
            dot += (double)p[3 * i + d] * (double)v;
        }
    }
    dot = chunk_sum(dot);
    if (threadIdx.x == 0 && cpart) cpart[c] = dot;
}

// r = f - q on free DOFs (0 on masked); chunk partials of f.f and r.r
// (all, when set: its active count is zeroed for k_breduce_init)
__global__ __launch_bounds__(BT) void k_binit(Chunks ch, const double* __restrict__ f, const double* __restrict__ q,
                                              const unsigned char* __restrict__ mask, double* __restrict__ r,
                                              double* __restrict__ cpart, int* __restrict__ all = nullptr) {
    if (all && blockIdx.x == 0 && threadIdx.x == 0) all[0] = 0;
    const int c = chunk_id(ch);
    double ff = 0.0, rr = 0.0;
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) {
        const bool fx = mask[k] != 0;
        const double fk = fx ? 0.0 : f[k];
        const double rk = fx ? 0.0 : fk - q[k];
        r[k] = rk;
        ff += fk * fk;
        rr += rk * rk;
    }
    ff = chunk_sum(ff);
    rr = chunk_sum(rr);
    if (threadIdx.x == 0) {
        cpart[2 * c] = ff;
        cpart[2 * c + 1] = rr;
    }
}

// per instance (one wave each): ||f||^2, r.r, stopping test at the initial guess; all[0]
// (zeroed by k_binit) counts the instances left running
__global__ __launch_bounds__(64) void k_breduce_init(int n_inst, const int* __restrict__ ico,
                                                     const double* __restrict__ cpart, double* __restrict__ S,
                                                     int* __restrict__ I, int* __restrict__ all, double tol) {
    const int i = blockIdx.x;
    double sm[2];
    wave_inst_sums<2>(ico, cpart, i, sm);
    if (threadIdx.x == 0) {
        const double ff = sm[0], rr = sm[1];
        double* s = S + (int64_t)i * NS;
        int* fl = I + (int64_t)i * NI;
        s[S_FF] = ff;
        s[S_RR] = rr;
        s[S_RZ] = rr;
        s[S_BETA] = 0.0;
        s[S_REL] = ff > 0.0 ? sqrt(rr / ff) : 0.0;
        fl[I_ITERS] = 0;
        fl[I_NOTPD] = 0;
        fl[I_ZERO_F] = ff == 0.0;
        fl[I_DONE] = (ff == 0.0 || rr <= tol * tol * ff) ? 1 : 0;
        if (!fl[I_DONE]) atomicAdd(all, 1);
    }
}

__global__ void k_bclear_flags(int n_inst, int* __restrict__ I) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) I[(int64_t)i * NI + I_DONE] = 0;
}

// u = 0 for instances without load (the reference's CG returns 0 for f = 0)
__global__ __launch_bounds__(BT) void k_bzero_u(Chunks ch, const int* __restrict__ I, double* __restrict__ u) {
    const int c = chunk_id(ch);
    if (!I[ch.inst[c] * NI + I_ZERO_F]) return;
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) u[k] = 0.0;
}

// per instance: alpha = rz / p.q, abort on p.q <= 0 (LinearStaticSolver.cu:404-407)
__global__ __launch_bounds__(64) void k_breduce_alpha(Insts in, const int* __restrict__ ico,
                                                      const double* __restrict__ cpart, double* __restrict__ S,
                                                      int* __restrict__ I) {
    const int i = inst_at(in, blockIdx.x);
    int* fl = I + (int64_t)i * NI;
    if (fl[I_DONE]) return;
    double sm[1];
    wave_inst_sums<1>(ico, cpart, i, sm);
    if (threadIdx.x == 0) {
        const double pq = sm[0];
        double* s = S + (int64_t)i * NS;
        s[S_PQ] = pq;
        if (pq <= 0.0) {
            fl[I_NOTPD] = 1;
            fl[I_DONE] = 1;
        } else {
            s[S_ALPHA] = s[S_RZ] / pq;
        }
    }
}

// p = z + beta p
__global__ __launch_bounds__(BT) void k_bdir(Chunks ch, double* __restrict__ p, const double* __restrict__ z,
                                             const double* __restrict__ S, const int* __restrict__ I) {
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    if (I[inst * NI + I_DONE]) return;
    const double beta = S[(int64_t)inst * NS + S_BETA];
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT)
        p[k] = beta != 0.0 ? z[k] + beta * p[k] : z[k];
}

// ---- single-reduction batched CG (Chronopoulos-Gear), the unpreconditioned solve --------
// Per iteration: the tangent applied to r, k_bgather_rw (w = K r per node and the chunk
// shares of r.w and r.r) and k_bcg_update, where every chunk sums its own instance's
// shares in chunk order, tests ||r|| / ||f|| < tol (the rule of the reduction kernels
// above) and forms beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
// before updating p, s = K p, u and r.  gamma_old / alpha_old sit in parity slots (the
// instance's leading chunk writes slot par, the others read par ^ 1); both start at
// (inf, 1), so the first iteration has beta = 0 and alpha = gamma / delta.

// w = K r (r on masked DOFs); chunk shares of r.w and r.r; block 0 clears the active count
__global__ __launch_bounds__(BT) void k_bgather_rw(Chunks ch, const int* __restrict__ g_off,
                                                   const int* __restrict__ g_idx, const double* __restrict__ part,
                                                   const unsigned char* __restrict__ mask,
                                                   const double* __restrict__ r, double* __restrict__ w,
                                                   double* __restrict__ cpart, const int* __restrict__ iflags,
                                                   int* __restrict__ all) {
    if (blockIdx.x == 0 && threadIdx.x == 0) all[0] = 0;
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    if (iflags[inst * NI + I_DONE]) {
        if (threadIdx.x == 0) cpart[2 * c] = cpart[2 * c + 1] = 0.0;
        return;
    }
    double rw = 0.0, rr = 0.0;
    for (int i = ch.beg[c] + threadIdx.x; i < ch.end[c]; i += BT) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        gather_slots3(g_idx, part, g_off[i], g_off[i + 1], s0, s1, s2);
        const double sv[3] = {s0, s1, s2};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double rk = r[3 * i + d];
            const double v = (mask && mask[3 * i + d]) ? rk : sv[d];
            w[3 * i + d] = v;
            rw += rk * v;
            rr += rk * rk;
        }
    }
    rw = chunk_sum(rw);
    rr = chunk_sum(rr);
    if (threadIdx.x == 0) {
        cpart[2 * c] = rw;
        cpart[2 * c + 1] = rr;
    }
}

__global__ __launch_bounds__(BT) void k_bcg_update(Chunks ch, const int* __restrict__ ico,
                                                   const double* __restrict__ cpart, double* __restrict__ u,
                                                   double* __restrict__ r, double* __restrict__ p,
                                                   double* __restrict__ sv, const double* __restrict__ w,
                                                   double* __restrict__ S, int* __restrict__ I,
                                                   int* __restrict__ all, double tol, int par) {
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    int* fl = I + (int64_t)inst * NI;
    if (fl[I_DONE]) return;
    // the instance's sums: strided over the block then reduced, the same order in every
    // chunk block of the instance
    double delta = 0.0, gamma = 0.0;
    for (int k = ico[inst] + threadIdx.x; k < ico[inst + 1]; k += BT) {
        delta += cpart[2 * k];
        gamma += cpart[2 * k + 1];
    }
    delta = chunk_sum(delta);
    gamma = chunk_sum(gamma);
    double* s = S + (int64_t)inst * NS;
    const bool lead = c == ico[inst] && threadIdx.x == 0;
    const double ff = s[S_FF];
    if (gamma < tol * tol * ff) {
        if (lead) {
            s[S_RR] = gamma;
            s[S_REL] = sqrt(gamma / ff);
            fl[I_DONE] = 1;
        }
        return;
    }
    const double beta = gamma / s[S_G0 + (par ^ 1)];
    const double den = delta - beta * gamma / s[S_A0 + (par ^ 1)];
    if (den <= 0.0) {  // p.Kp <= 0 (LinearStaticSolver.cu:404-407)
        if (lead) {
            s[S_PQ] = den;
            fl[I_NOTPD] = 1;
            fl[I_DONE] = 1;
        }
        return;
    }
    const double alpha = gamma / den;
    if (lead) {
        s[S_G0 + par] = gamma;
        s[S_A0 + par] = alpha;
        s[S_RR] = gamma;
        s[S_REL] = sqrt(gamma / ff);
        s[S_ALPHA] = alpha;
        s[S_BETA] = beta;
        s[S_PQ] = den;
        fl[I_ITERS] += 1;
        atomicAdd(all, 1);
    }
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) {
        const double rk = r[k];
        const double pk = rk + beta * p[k];
        const double sk = w[k] + beta * sv[k];
        p[k] = pk;
        sv[k] = sk;
        u[k] += alpha * pk;
        r[k] = rk - alpha * sk;
    }
}

// chunk partials of a.c over every DOF of the chunk
__global__ __launch_bounds__(BT) void k_bdot_ac(Chunks ch, const double* __restrict__ a, const double* __restrict__ c,
                                                double* __restrict__ cpart) {
    const int k = chunk_id(ch);
    double s = 0.0;
    for (int64_t j = 3 * (int64_t)ch.beg[k] + threadIdx.x; j < 3 * (int64_t)ch.end[k]; j += BT) s += a[j] * c[j];
    s = chunk_sum(s);
    if (threadIdx.x == 0) cpart[k] = s;
}

// per instance (one wave each): the sum of its chunk partials
__global__ __launch_bounds__(64) void k_binst_sum(int n_inst, const int* __restrict__ ico,
                                                  const double* __restrict__ cpart, double* __restrict__ out) {
    double sm[1];
    wave_inst_sums<1>(ico, cpart, blockIdx.x, sm);
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// parity slots of a fresh solve
__global__ void k_bcg_init_slots(int n_inst, double* __restrict__ S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    double* s = S + (int64_t)i * NS;
    s[S_G0] = s[S_G1] = INFINITY;
    s[S_A0] = s[S_A1] = 1.0;
}

// chunk partials of r.r (the test after a solve that ran out of iterations)
__global__ __launch_bounds__(BT) void k_bdot_rr(Chunks ch, const double* __restrict__ r, double* __restrict__ cpart) {
    const int c = chunk_id(ch);
    double rr = 0.0;
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) rr += r[k] * r[k];
    rr = chunk_sum(rr);
    if (threadIdx.x == 0) cpart[c] = rr;
}

__global__ __launch_bounds__(64) void k_breduce_tail(int n_inst, const int* __restrict__ ico,
                                                     const double* __restrict__ cpart, double* __restrict__ S,
                                                     int* __restrict__ I, double tol) {
    const int i = blockIdx.x;
    int* fl = I + (int64_t)i * NI;
    if (fl[I_DONE]) return;
    double sm[1];
    wave_inst_sums<1>(ico, cpart, i, sm);
    if (threadIdx.x == 0) {
        const double rr = sm[0];
        double* s = S + (int64_t)i * NS;
        s[S_RR] = rr;
        s[S_REL] = sqrt(rr / s[S_FF]);
        if (rr < tol * tol * s[S_FF]) fl[I_DONE] = 1;
    }
}

// ---- FP32 Jacobi-Chebyshev preconditioner ----------------------------------------------
// z ~= K^-1 r by `degree` Chebyshev steps on D^-1 K over [lmax/ratio, lmax] (Saad,
// Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1), all in FP32; the
// polynomial is positive on (0, lmax], so the preconditioner is SPD up to rounding and
// the outer FP64 CG uses the flexible (Polak-Ribiere) beta.

// rho_k of the three-term recursion: rho_0 = 1/sigma, rho_k = 1/(2 sigma - rho_{k-1})
__device__ __forceinline__ float cheb_rho(float sigma, int k) {
    float rho = 1.0f / sigma;
    for (int j = 0; j < k; j++) rho = 1.0f / (2.0f * sigma - rho);
    return rho;
}

// res = r, d = D^-1 res / theta, z = d
__global__ __launch_bounds__(BT) void k_cheb_init(Chunks ch, const double* __restrict__ r,
                                                  const float* __restrict__ dinv, const double* __restrict__ C,
                                                  const int* __restrict__ I, float* __restrict__ res,
                                                  float* __restrict__ d, float* __restrict__ zc) {
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    if (I[inst * NI + I_DONE]) return;
    const float ith = (float)(1.0 / C[(int64_t)inst * NC + C_THETA]);
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) {
        const float rk = (float)r[k];
        const float dk = dinv[k] * rk * ith;
        res[k] = rk;
        d[k] = dk;
        zc[k] = dk;
    }
}


// The FP32 operator's gather fused with the Chebyshev step: q = the node's partial sums (d
// on masked DOFs, the operator's identity rows), then res -= q, d = a d + b D^-1 res,
// zc += d, as k_bgather<float> followed by k_cheb_step.
__global__ __launch_bounds__(BT) void k_bgather_cheb(Chunks ch, const int* __restrict__ g_off,
                                                     const int* __restrict__ g_idx, const float* __restrict__ part,
                                                     const unsigned char* __restrict__ mask, int k,
                                                     const float* __restrict__ dinv, const double* __restrict__ C,
                                                     const int* __restrict__ I, float* __restrict__ res,
                                                     float* __restrict__ d, float* __restrict__ zc) {
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    if (I[inst * NI + I_DONE]) return;
    const float sigma = (float)C[(int64_t)inst * NC + C_SIGMA];
    const float delta = (float)C[(int64_t)inst * NC + C_DELTA];
    const float rp = cheb_rho(sigma, k - 1), rk = cheb_rho(sigma, k);
    const float a = rk * rp, bcoef = 2.0f * rk / delta;
    for (int i = ch.beg[c] + threadIdx.x; i < ch.end[c]; i += BT) {
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
        gather_slots3(g_idx, part, g_off[i], g_off[i + 1], s0, s1, s2);
        const float sv[3] = {s0, s1, s2};
#pragma unroll
        for (int e = 0; e < 3; e++) {
            const int64_t j = 3 * (int64_t)i + e;
            const float dold = d[j];
            const float qj = mask[j] ? dold : sv[e];
            const float rj = res[j] - qj;
            res[j] = rj;
            const float dj = a * dold + bcoef * dinv[j] * rj;
            d[j] = dj;
            zc[j] += dj;
        }
    }
}

// In-plane fields of a node plane with normal ax at the node's scaled offset r (coarse modes
// 6..10): with (s, t) the offset's in-plane coordinates along (e_s, e_t), the two lowest
// in-plane axes,
//     6: s t e_s + (t^2 - s^2)/2 e_t,  7: (s^2 - t^2)/2 e_s + s t e_t   (bending warping)
//     8: s e_s + t e_t,  9: t e_s + s e_t,  10: s e_s - t e_t            (stretch, shears)
// Returns the contribution sum_j c[j] mode_j(r) . e_d for component d (c[0..4] = modes 6..10).
__device__ __forceinline__ double inplane_value(const double* c, const double* r, int ax, int d) {
    const int es = ax == 0 ? 1 : 0, et = ax == 2 ? 1 : 2;
    const double s = r[es], t = r[et];
    if (d == es) return c[0] * s * t + c[1] * 0.5 * (s * s - t * t) + c[2] * s + c[3] * t + c[4] * s;
    if (d == et) return c[0] * 0.5 * (t * t - s * s) + c[1] * s * t + c[2] * t + c[3] * s - c[4] * t;
    return 0.0;
}

// The coarse correction's prolongation (node_plane non-null): (P x)_k = t_p + theta_p x r_n
// (+ the in-plane fields with cm = 11) at DOF k = 3 n + d of a free node n of plane p (0 on
// fixed nodes); x holds cm values per plane.
struct CoarseProlong {
    const int* node_plane;
    const double* node_r;
    const double* x;
    int cm = 6;
    const int* plane_axis = nullptr;
};
__device__ __forceinline__ double coarse_value(const CoarseProlong& cp, int64_t k) {
    const int64_t n = k / 3;
    const int d = (int)(k - 3 * n);
    const int pl = cp.node_plane[n];
    if (pl < 0) return 0.0;
    const double* t = cp.x + cp.cm * (int64_t)pl;
    const double* r = cp.node_r + 3 * n;
    double v;
    if (d == 0)
        v = t[0] + t[4] * r[2] - t[5] * r[1];
    else if (d == 1)
        v = t[1] + t[5] * r[0] - t[3] * r[2];
    else
        v = t[2] + t[3] * r[1] - t[4] * r[0];
    if (cp.cm > 6) v += inplane_value(t + 6, r, cp.plane_axis[pl], d);
    return v;
}

// z = P x on every DOF (FP64)
__global__ void k_coarse_prolong64(int64_t n3, CoarseProlong cp, double* __restrict__ z) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n3) z[k] = coarse_value(cp, k);
}

// z = z32 in FP64 (plus P x of the coarse solve when set up: the additive two-level
// preconditioner); chunk partials of z.r and z.r_old
__global__ __launch_bounds__(BT) void k_cheb_fin(Chunks ch, const float* __restrict__ zc,
                                                 const double* __restrict__ r, const double* __restrict__ rold,
                                                 const int* __restrict__ I, double* __restrict__ z,
                                                 double* __restrict__ cpart, CoarseProlong cp) {
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    double zr = 0.0, zro = 0.0;
    if (!I[inst * NI + I_DONE]) {
        for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) {
            const double zk = (double)zc[k] + (cp.node_plane ? coarse_value(cp, k) : 0.0);
            z[k] = zk;
            zr += zk * r[k];
            zro += zk * rold[k];
        }
    }
    zr = chunk_sum(zr);
    zro = chunk_sum(zro);
    if (threadIdx.x == 0) {
        cpart[2 * c] = zr;
        cpart[2 * c + 1] = zro;
    }
}

// per instance after the first preconditioning: rz = z.r, p = z (beta = 0)
__global__ __launch_bounds__(64) void k_breduce_rz0(Insts in, const int* __restrict__ ico,
                                                    const double* __restrict__ cpart, double* __restrict__ S,
                                                    const int* __restrict__ I) {
    const int i = inst_at(in, blockIdx.x);
    if (I[(int64_t)i * NI + I_DONE]) return;
    double sm[2];
    wave_inst_sums<2>(ico, cpart, i, sm);
    if (threadIdx.x == 0) {
        S[(int64_t)i * NS + S_RZ] = sm[0];
        S[(int64_t)i * NS + S_BETA] = 0.0;
    }
}

// u += alpha p, r_old = r, r -= alpha q; chunk partial of r.r
// (all[0] zeroed for k_breduce_check's count)
__global__ __launch_bounds__(BT) void k_bupdate_keep(Chunks ch, double* __restrict__ u, double* __restrict__ r,
                                                     double* __restrict__ rold, const double* __restrict__ p,
                                                     const double* __restrict__ q, const double* __restrict__ S,
                                                     const int* __restrict__ I, double* __restrict__ cpart,
                                                     int* __restrict__ all) {
    if (blockIdx.x == 0 && threadIdx.x == 0) all[0] = 0;
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    if (I[inst * NI + I_DONE]) {
        if (threadIdx.x == 0) cpart[c] = 0.0;
        return;
    }
    const double alpha = S[(int64_t)inst * NS + S_ALPHA];
    double rr = 0.0;
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) {
        u[k] += alpha * p[k];
        const double r0 = r[k];
        rold[k] = r0;
        const double rk = r0 - alpha * q[k];
        r[k] = rk;
        rr += rk * rk;
    }
    rr = chunk_sum(rr);
    if (threadIdx.x == 0) cpart[c] = rr;
}

// per instance: stopping test only (preconditioned path); all[0] (zeroed by
// k_bupdate_keep) counts the instances left running
__global__ __launch_bounds__(64) void k_breduce_check(Insts in, const int* __restrict__ ico,
                                                      const double* __restrict__ cpart, double* __restrict__ S,
                                                      int* __restrict__ I, int* __restrict__ all, double tol) {
    const int i = inst_at(in, blockIdx.x);
    int* fl = I + (int64_t)i * NI;
    if (fl[I_DONE]) return;
    double sm[1];
    wave_inst_sums<1>(ico, cpart, i, sm);
    if (threadIdx.x == 0) {
        const double rr = sm[0];
        double* s = S + (int64_t)i * NS;
        fl[I_ITERS] += 1;
        s[S_RR] = rr;
        s[S_REL] = sqrt(rr / s[S_FF]);
        if (rr < tol * tol * s[S_FF])
            fl[I_DONE] = 1;
        else
            atomicAdd(all, 1);
    }
}

// per instance: flexible beta = z.(r - r_old) / rz_old
__global__ __launch_bounds__(64) void k_breduce_fbeta(Insts in, const int* __restrict__ ico,
                                                      const double* __restrict__ cpart, double* __restrict__ S,
                                                      const int* __restrict__ I) {
    const int i = inst_at(in, blockIdx.x);
    if (I[(int64_t)i * NI + I_DONE]) return;
    double sm[2];
    wave_inst_sums<2>(ico, cpart, i, sm);
    if (threadIdx.x == 0) {
        double* s = S + (int64_t)i * NS;
        s[S_BETA] = (sm[0] - sm[1]) / s[S_RZ];
        s[S_RZ] = sm[0];
    }
}

// ---- power iteration for lambda_max(D^-1 K) per instance (FP32) -----------------------
__global__ __launch_bounds__(BT) void k_pow_start(Chunks ch, const unsigned char* __restrict__ mask,
                                                  float* __restrict__ v) {
    const int c = chunk_id(ch);
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) {
        // deterministic pseudo-random start in [0.5, 1.5)
        uint32_t h = (uint32_t)k * 2654435761u;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        v[k] = mask[k] ? 0.0f : 0.5f + (float)(h & 0xffffu) / 65536.0f;
    }
}

// y = D^-1 w on free DOFs; chunk partial of |y|^2
__global__ __launch_bounds__(BT) void k_pow_scale(Chunks ch, const float* __restrict__ w,
                                                  const float* __restrict__ dinv,
                                                  const unsigned char* __restrict__ mask, float* __restrict__ y,
                                                  double* __restrict__ cpart) {
    const int c = chunk_id(ch);
    double yy = 0.0;
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) {
        const float yk = mask[k] ? 0.0f : dinv[k] * w[k];
        y[k] = yk;
        yy += (double)yk * yk;
    }
    yy = chunk_sum(yy);
    if (threadIdx.x == 0) cpart[c] = yy;
}

// per instance: lambda = |y| / |v_prev| (|v_prev| = 1 after the first pass), stored in C
__global__ __launch_bounds__(64) void k_pow_reduce(int n_inst, const int* __restrict__ ico,
                                                   const double* __restrict__ cpart, double* __restrict__ C) {
    double sm[1];
    wave_inst_sums<1>(ico, cpart, blockIdx.x, sm);
    if (threadIdx.x == 0) C[(int64_t)blockIdx.x * NC + C_LMAX] = sqrt(sm[0]);
}

// v = y / |y| per instance (the previous pass's norm)
__global__ __launch_bounds__(BT) void k_pow_norm(Chunks ch, const float* __restrict__ y, const double* __restrict__ C,
                                                 float* __restrict__ v) {
    const int c = chunk_id(ch);
    const double nrm = C[(int64_t)ch.inst[c] * NC + C_LMAX];
    const float s = nrm > 0.0 ? (float)(1.0 / nrm) : 0.0f;
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) v[k] = y[k] * s;
}

// Chebyshev interval [b/ratio, b] with b = safety * lambda estimate
__global__ void k_cheb_params(int n_inst, double safety, double ratio, double* __restrict__ C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    double* c = C + (int64_t)i * NC;
    const double bmax = safety * c[C_LMAX];
    const double amin = bmax / ratio;
    c[C_THETA] = 0.5 * (bmax + amin);
    c[C_DELTA] = 0.5 * (bmax - amin);
    c[C_SIGMA] = c[C_THETA] / c[C_DELTA];
}

// D^-1 in FP32 from the gathered diagonal (1 on fixed DOFs, where the gather returns p = 1)
__global__ void k_inv_diag(int64_t n, const double* __restrict__ dg, float* __restrict__ dinv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dinv[i] = dg[i] > 0.0 ? (float)(1.0 / dg[i]) : 1.0f;
}

__global__ void k_fill(int64_t n, double v, double* __restrict__ a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_elem_inst(int n_inst, const int* __restrict__ elem_off, int* __restrict__ elem_inst) {
    const int i = blockIdx.x;
    if (i >= n_inst) return;
    for (int e = elem_off[i] + threadIdx.x; e < elem_off[i + 1]; e += blockDim.x) elem_inst[e] = i;
}

// ---- coarse-space geometry on the device (the host sorts one coordinate per node) -------
// min and max of X, Y, Z over each instance's nodes, one workgroup of 256 per instance:
// bbox[6 i ..] = (lo_x, lo_y, lo_z, hi_x, hi_y, hi_z)
__global__ __launch_bounds__(256) void k_geom_bbox(const int* __restrict__ node_off, const double* __restrict__ X,
                                                   const double* __restrict__ Y, const double* __restrict__ Z,
                                                   double* __restrict__ bbox) {
    const int i = blockIdx.x;
    const double* c[3] = {X, Y, Z};
    double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int n = node_off[i] + (int)threadIdx.x; n < node_off[i + 1]; n += 256)
        for (int d = 0; d < 3; d++) {
            v[d] = fmin(v[d], c[d][n]);
            v[3 + d] = fmax(v[3 + d], c[d][n]);
        }
    __shared__ double red[4][6];
    for (int k = 0; k < 6; k++) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_xor(x, o, 64);
            x = k < 3 ? fmin(x, y) : fmax(x, y);
        }
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double x = red[0][k];
        for (int w = 1; w < 4; w++) x = k < 3 ? fmin(x, red[w][k]) : fmax(x, red[w][k]);
        bbox[6 * (int64_t)i + k] = x;
    }
}

// per node: its coordinate along its instance's long axis, and 1 where a DOF is fixed
__global__ void k_geom_axis(int n_nodes, const int* __restrict__ node_inst, const int* __restrict__ axis,
                            const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
                            const unsigned char* __restrict__ mask, double* __restrict__ cax,
                            unsigned char* __restrict__ fixed) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    const int a = axis[node_inst[n]];
    cax[n] = a == 0 ? X[n] : (a == 1 ? Y[n] : Z[n]);
    fixed[n] = (mask[3 * (int64_t)n] | mask[3 * (int64_t)n + 1] | mask[3 * (int64_t)n + 2]) ? 1 : 0;
}

// The per-instance maximum of a per-element value v (>= 0): a wave whose elements all
// belong to one instance reduces first and adds one atomic (a large instance would
// otherwise serialise all its elements' atomics on one address); other waves, at an
// instance boundary or the end, take one atomic per element.
template <typename T>
__device__ __forceinline__ void inst_max(T* __restrict__ out, int inst, T v) {
    const int i0 = __shfl(inst, 0, 64);
    if (__all(inst == i0)) {
        for (int o = 32; o > 0; o >>= 1) {
            const T w = __shfl_xor(v, o, 64);
            v = w > v ? w : v;
        }
        if ((threadIdx.x & 63) == 0) atomicMax(&out[i0], v);
    } else if (inst >= 0) {
        atomicMax(&out[inst], v);
    }
}

// per element: its extent along its instance's axis; the maximum per instance by an atomic
// max of the bit patterns (non-negative doubles order as their patterns)
__global__ __launch_bounds__(256) void k_geom_elem_extent(int n_elem, int nn, const int* __restrict__ conn,
                                                          const int* __restrict__ elem_inst,
                                                          const double* __restrict__ cax,
                                                          unsigned long long* __restrict__ emax) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long bits = 0;
    int inst = -1;
    if (e < n_elem) {
        double lo = 1e300, hi = -1e300;
        for (int a = 0; a < nn; a++) {
            const double v = cax[conn[e + (int64_t)a * n_elem]];
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        bits = (unsigned long long)__double_as_longlong(fmax(hi - lo, 0.0));
        inst = elem_inst[e];
    }
    inst_max(emax, inst, bits);
}

// per element: the span of its nodes' local planes; the maximum per instance
__global__ __launch_bounds__(256) void k_geom_elem_span(int n_elem, int nn, const int* __restrict__ conn,
                                                        const int* __restrict__ elem_inst,
                                                        const int* __restrict__ local, int* __restrict__ span) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int d = 0, inst = -1;
    if (e < n_elem) {
        int lo = 1 << 30, hi = -1;
        for (int a = 0; a < nn; a++) {
            const int q = local[conn[e + (int64_t)a * n_elem]];
            lo = min(lo, q);
            hi = max(hi, q);
        }
        d = hi - lo;
        inst = elem_inst[e];
    }
    inst_max(span, inst, d);
}

// per plane, one workgroup of 256: the centroid c of its free nodes and their largest
// distance from it (cen[3 p ..], size[p]; zero for a plane without free nodes)
__global__ __launch_bounds__(256) void k_geom_plane_frame(const int* __restrict__ plane_off,
                                                          const int* __restrict__ plane_nodes,
                                                          const double* __restrict__ X, const double* __restrict__ Y,
                                                          const double* __restrict__ Z, double* __restrict__ cen,
                                                          double* __restrict__ size) {
    const int p = blockIdx.x;
    const int k0 = plane_off[p], k1 = plane_off[p + 1];
    __shared__ double red[4][3];
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 256) {
        const int n = plane_nodes[k];
        s[0] += X[n];
        s[1] += Y[n];
        s[2] += Z[n];
    }
    for (int d = 0; d < 3; d++) {
        for (int o = 32; o > 0; o >>= 1) s[d] += __shfl_xor(s[d], o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][d] = s[d];
    }
    __syncthreads();
    double c[3];
    for (int d = 0; d < 3; d++) c[d] = k1 > k0 ? (red[0][d] + red[1][d] + red[2][d] + red[3][d]) / (k1 - k0) : 0.0;
    __syncthreads();
    double r2 = 0.0;
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 256) {
        const int n = plane_nodes[k];
        const double a = X[n] - c[0], b = Y[n] - c[1], e = Z[n] - c[2];
        r2 = fmax(r2, a * a + b * b + e * e);
    }
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][0] = r2;
    __syncthreads();
    if (threadIdx.x == 0) {
        cen[3 * (int64_t)p] = c[0];
        cen[3 * (int64_t)p + 1] = c[1];
        cen[3 * (int64_t)p + 2] = c[2];
        size[p] = sqrt(fmax(fmax(red[0][0], red[1][0]), fmax(red[2][0], red[3][0])));
    }
}

// per node: r = (X - c) / size of its plane (zero on fixed nodes)
__global__ void k_geom_node_r(int n_nodes, const int* __restrict__ node_plane, const double* __restrict__ X,
                              const double* __restrict__ Y, const double* __restrict__ Z,
                              const double* __restrict__ cen, const double* __restrict__ size,
                              double* __restrict__ node_r) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    const int g = node_plane[n];
    double r[3] = {0.0, 0.0, 0.0};
    if (g >= 0) {
        const double sc = size[g] > 0.0 ? 1.0 / size[g] : 1.0;
        r[0] = (X[n] - cen[3 * (int64_t)g]) * sc;
        r[1] = (Y[n] - cen[3 * (int64_t)g + 1]) * sc;
        r[2] = (Z[n] - cen[3 * (int64_t)g + 2]) * sc;
    }
    node_r[3 * (int64_t)n] = r[0];
    node_r[3 * (int64_t)n + 1] = r[1];
    node_r[3 * (int64_t)n + 2] = r[2];
}

__global__ void k_d2f(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}


// ---- two-level coarse correction --------------------------------------------------------
// Coarse space: for each plane p of nodes across an instance's long axis, the 6 rigid
// motions u_n = t_p + theta_p x r_n of its free nodes (r_n = (X_n - c_p) / size_p), so
// P^T v = [sum_n v_n ; sum_n r_n x v_n] per plane, and with cm = 11 the five in-plane
// fields of inplane_value.  Elements span at most W consecutive planes, so P^T K P couples
// planes at most W apart: grouping W planes per block makes it block-tridiagonal (blocks
// of NB = cm W).  It is assembled by probing with the batched operator (planes coloured
// mod 2W+1), factored as D_k = A_kk - G_k A_(k-1,k)^T with G_k = A_(k,k-1) D_(k-1)^-1, and
// solved by a block forward and backward sweep.
constexpr int CNB_MAX = 22;  // 11 modes with W <= 2, 6 with W = 3

// mode j < cm of a plane with normal ax at the scaled offset r
__device__ __forceinline__ void coarse_mode(int j, const double* r, int ax, double v[3]) {
    v[0] = v[1] = v[2] = 0.0;
    if (j < 3) {
        v[j] = 1.0;
    } else if (j < 6) {  // e_(j-3) x r
        const int a = j - 3;
        const double e[3] = {a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0};
        v[0] = e[1] * r[2] - e[2] * r[1];
        v[1] = e[2] * r[0] - e[0] * r[2];
        v[2] = e[0] * r[1] - e[1] * r[0];
    } else {
        double c[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        c[j - 6] = 1.0;
        for (int d = 0; d < 3; d++) v[d] = inplane_value(c, r, ax, d);
    }
}

// q = P e_(p, j) summed over the planes p with local index = color mod (2W + 1)
__global__ void k_coarse_probe_vec(int n_nodes, const int* __restrict__ node_plane, const double* __restrict__ node_r,
                                   const int* __restrict__ plane_local, const int* __restrict__ plane_axis, int period,
                                   int color, int j, double* __restrict__ q) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    const int pl = node_plane[n];
    double v[3] = {0.0, 0.0, 0.0};
    if (pl >= 0 && plane_local[pl] % period == color) coarse_mode(j, node_r + 3 * (int64_t)n, plane_axis[pl], v);
    q[3 * (int64_t)n] = v[0];
    q[3 * (int64_t)n + 1] = v[1];
    q[3 * (int64_t)n + 2] = v[2];
}

// the cm sums of P^T v over plane p (normal ax), in lane order then by wave reduction (64
// threads) (NT threads: a workgroup of NT; with NT > 64 the waves' sums combine through LDS
// and every thread returns the total); out[cm..CM_MAX) are zero
template <int NT = 64>
__device__ __forceinline__ void plane_restrict(int p, const int* __restrict__ plane_off,
                                               const int* __restrict__ plane_nodes, const double* __restrict__ node_r,
                                               const double* __restrict__ v, int cm, int ax, double out[CM_MAX]) {
    double acc[CM_MAX];
#pragma unroll
    for (int i = 0; i < CM_MAX; i++) acc[i] = 0.0;
    const int es = ax == 0 ? 1 : 0, et = ax == 2 ? 1 : 2;
    for (int k = plane_off[p] + (int)threadIdx.x; k < plane_off[p + 1]; k += NT) {
        const int n = plane_nodes[k];
        const double* r = node_r + 3 * (int64_t)n;
        const double a = v[3 * (int64_t)n], b = v[3 * (int64_t)n + 1], c = v[3 * (int64_t)n + 2];
        acc[0] += a;
        acc[1] += b;
        acc[2] += c;
        acc[3] += r[1] * c - r[2] * b;
        acc[4] += r[2] * a - r[0] * c;
        acc[5] += r[0] * b - r[1] * a;
        if (cm > 6) {
            const double s = r[es], t = r[et];
            const double vs = v[3 * (int64_t)n + es], vt = v[3 * (int64_t)n + et];
            acc[6] += s * t * vs + 0.5 * (t * t - s * s) * vt;
            acc[7] += 0.5 * (s * s - t * t) * vs + s * t * vt;
            acc[8] += s * vs + t * vt;
            acc[9] += t * vs + s * vt;
            acc[10] += s * vs - t * vt;
        }
    }
#pragma unroll
    for (int i = 0; i < CM_MAX; i++) {
        double x = acc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        out[i] = x;
    }
    if constexpr (NT > 64) {
        __shared__ double red[NT / 64][CM_MAX];
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int i = 0; i < CM_MAX; i++) red[threadIdx.x >> 6][i] = out[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < CM_MAX; i++) {
            double x = 0.0;
#pragma unroll
            for (int w = 0; w < NT / 64; w++) x += red[w][i];
            out[i] = x;
        }
    }
}

// A(p', (p*, j)) = (P^T w)_p' for the coloured plane p* within W of p' (same instance);
// diagonal blocks and the lower blocks A_(k, k-1) are stored (row-major NB x NB).
template <int NT>
__global__ __launch_bounds__(NT) void k_coarse_probe_store(int W, int cm, int period, int color, int j,
                                                           const int* __restrict__ plane_off,
                                                           const int* __restrict__ plane_nodes,
                                                           const int* __restrict__ plane_local,
                                                           const int* __restrict__ plane_inst,
                                                           const int* __restrict__ plane_axis,
                                                           const int* __restrict__ inst_plane_n,
                                                           const double* __restrict__ node_r,
                                                           const double* __restrict__ w, double* __restrict__ cdiag,
                                                           double* __restrict__ clow) {
    const int pp = blockIdx.x;  // p'
    const int lp = plane_local[pp];
    // the plane with local index = color (mod period) within W of p'
    const int dm = ((lp - color) % period + period) % period;
    int ls = -1;
    if (dm <= W)
        ls = lp - dm;
    else if (period - dm <= W)
        ls = lp + (period - dm);
    if (ls < 0 || ls >= inst_plane_n[plane_inst[pp]]) return;
    const int ps = pp + (ls - lp);
    double out[CM_MAX];
    plane_restrict<NT>(pp, plane_off, plane_nodes, node_r, w, cm, plane_axis[pp], out);
    if (threadIdx.x != 0) return;
    const int NB = cm * W;
    const int bp = pp / W, bs = ps / W;
    const int row0 = cm * (pp % W), col = cm * (ps % W) + j;
    double* dst = bs == bp ? cdiag + (int64_t)bp * NB * NB : (bs == bp - 1 ? clow + (int64_t)bp * NB * NB : nullptr);
    if (!dst) return;  // the upper block is the transpose of the next lower one
    for (int i = 0; i < cm; i++) dst[(row0 + i) * NB + col] = out[i];
}

// Block factorisation per instance, one wave: cdiag[k] becomes D_k^-1, cg[k] = G_k, with
// D_k = A_kk - G_k A_(k-1,k) and G_k = A_(k,k-1) D_(k-1)^-1.  The chain over k is serial, so
// a step is kept short: lane r holds row r of the step's blocks in registers (NB a
// compile-time constant), the products read the other operand from LDS as broadcasts, and
// the Cholesky factorisation of D_k runs in registers, the columns passed between lanes by
// shuffles (no barrier inside it), each pivot's reciprocal formed once so that no division
// sits on the serial path.  A pivot at or below 1e-10 of its diagonal entry drops that
// coarse DOF (its row and column of the inverse are zero: no rigid mode of a plane without
// the free nodes to carry it).  Lane c then solves L y = e_c, L^T x = y for column c of
// D_k^-1.
template <int NB>
__global__ __launch_bounds__(64) void k_coarse_factor(const int* __restrict__ inst_cblk,
                                                      double* __restrict__ cdiag, const double* __restrict__ clow,
                                                      double* __restrict__ cg) {
    static_assert(NB <= 12, "rows of at most 12 doubles per lane (larger blocks: k_coarse_factor_lds)");
    constexpr int nn = NB * NB;
    __shared__ double sDinv[nn], sA[nn], sM[nn], sL[nn];
    const int i = blockIdx.x;
    const int k0 = inst_cblk[i], k1 = inst_cblk[i + 1];
    const int lane = (int)threadIdx.x;
    const int r = lane < NB ? lane : NB - 1;  // idle lanes shadow the last row, store nothing
    // row r of the next step's diagonal and lower blocks, fetched a step ahead
    double nD[NB], nA[NB];
    auto fetch = [&](int k) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
            nD[c] = cdiag[(int64_t)k * nn + r * NB + c];
            nA[c] = k > k0 ? clow[(int64_t)k * nn + r * NB + c] : 0.0;  // A_(k, k-1)
        }
    };
    if (k0 < k1) fetch(k0);
    for (int k = k0; k < k1; k++) {
        double m[NB], a[NB];
#pragma unroll
        for (int c = 0; c < NB; c++) {
            m[c] = nD[c];
            a[c] = nA[c];
        }
        if (k + 1 < k1) fetch(k + 1);
        double* G = cg + (int64_t)k * nn;
        if (k > k0) {
            if (lane < NB)
#pragma unroll
                for (int c = 0; c < NB; c++) sA[r * NB + c] = a[c];
            __syncthreads();
            // row r of G = A Dprev^-1 (Dprev^-1 read row by row), then of M -= G A^T
            double g[NB];
#pragma unroll
            for (int c = 0; c < NB; c++) g[c] = 0.0;
#pragma unroll
            for (int t = 0; t < NB; t++)
#pragma unroll
                for (int c = 0; c < NB; c++) g[c] += a[t] * sDinv[t * NB + c];
            if (lane < NB)
#pragma unroll
                for (int c = 0; c < NB; c++) G[r * NB + c] = g[c];
#pragma unroll
            for (int c = 0; c < NB; c++) {
                double v = 0.0;
#pragma unroll
                for (int t = 0; t < NB; t++) v += g[t] * sA[c * NB + t];
                m[c] -= v;
            }
        } else if (lane < NB) {
#pragma unroll
            for (int c = 0; c < NB; c++) G[r * NB + c] = 0.0;
        }
        // symmetrise (the probed entries agree to rounding): the lower triangle takes the
        // mean of the two
        if (lane < NB)
#pragma unroll
            for (int c = 0; c < NB; c++) sM[r * NB + c] = m[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NB; c++)
            if (c < r) m[c] = 0.5 * (sM[c * NB + r] + m[c]);
        // Cholesky in registers: after column c, m[c] of lane r > c holds L[r][c]
        double dg = m[0];
#pragma unroll
        for (int c = 1; c < NB; c++) dg = lane == c ? m[c] : dg;
        bool dead[NB];
        double rl[NB];  // 1 / L[c][c] (0 on dead columns)
#pragma unroll
        for (int c = 0; c < NB; c++) {
            const double d = __shfl(m[c], c, 64), d0 = __shfl(dg, c, 64);
            const bool dd = !(d > 1e-10 * fabs(d0)) || d0 <= 0.0;
            dead[c] = dd;
            const double lc = dd ? 0.0 : sqrt(d);
            rl[c] = dd ? 0.0 : 1.0 / lc;
            m[c] = r > c ? m[c] * rl[c] : (r == c ? lc : m[c]);
#pragma unroll
            for (int j = c + 1; j < NB; j++) {
                const double ljc = __shfl(m[c], j, 64);
                if (r >= j) m[j] -= m[c] * ljc;
            }
        }
        if (lane < NB)
#pragma unroll
            for (int c = 0; c < NB; c++)
                if (c <= r) sL[r * NB + c] = m[c];
        __syncthreads();
        // column c of D_k^-1: L y = e_c, L^T x = y (dead rows skipped)
        double* D = cdiag + (int64_t)k * nn;
        if (lane < NB) {
            const int c = lane;
            double y[NB], x[NB];
#pragma unroll
            for (int q = 0; q < NB; q++) {
                double v = (q == c) ? 1.0 : 0.0;
#pragma unroll
                for (int t = 0; t < q; t++) v -= sL[q * NB + t] * y[t];
                y[q] = v * rl[q];
            }
#pragma unroll
            for (int q = NB - 1; q >= 0; q--) {
                double v = y[q];
#pragma unroll
                for (int t = q + 1; t < NB; t++) v -= sL[t * NB + q] * x[t];
                x[q] = v * rl[q];
            }
#pragma unroll
            for (int q = 0; q < NB; q++) {
                sDinv[q * NB + c] = x[q];
                D[q * NB + c] = x[q];
            }
        }
        __syncthreads();
    }
}

// In-place inverse of the SPD NB x NB matrix m (LDS) by Cholesky: the workgroup factors
// together, then lanes 0..NB-1 solve one column each.  A pivot at or below 1e-10 of its diagonal entry
// drops that coarse DOF (its row and column of the inverse are zero: no rigid mode of a
// plane without the free nodes to carry it).  Called by all threads of the workgroup.
template <int NB>
__device__ void spd_inverse(double* m, double* L, int* dead) {
    // right-looking Cholesky, the workgroup's threads over each step's column and trailing
    // block: L starts as the lower triangle of m and is updated in place
    __shared__ double d0s[CNB_MAX];
    const int t = threadIdx.x, nt = blockDim.x;
    for (int e = t; e < NB * NB; e += nt) L[e] = (e % NB) <= (e / NB) ? m[e] : 0.0;
    if (t < NB) d0s[t] = m[t * NB + t];
    __syncthreads();
    for (int c = 0; c < NB; c++) {
        const double d = L[c * NB + c], d0 = d0s[c];
        const bool dd = !(d > 1e-10 * fabs(d0)) || d0 <= 0.0;
        const double lc = dd ? 0.0 : sqrt(d);
        __syncthreads();  // every thread has read the pivot
        if (t == 0) {
            dead[c] = dd ? 1 : 0;
            L[c * NB + c] = lc;
        }
        for (int i = c + 1 + t; i < NB; i += nt) L[i * NB + c] = dd ? 0.0 : L[i * NB + c] / lc;
        __syncthreads();
        const int m2 = NB - c - 1;
        for (int e = t; e < m2 * m2; e += nt) {
            const int i = c + 1 + e / m2, j = c + 1 + e % m2;
            if (j <= i) L[i * NB + j] -= L[i * NB + c] * L[j * NB + c];
        }
        __syncthreads();
    }
    // column c of the inverse: L y = e_c, L^T x = y (dead rows skipped)
    const int c = threadIdx.x;
    if (c < NB) {
        // NB a compile-time constant: y and x stay in registers
        double y[NB], x[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            double v = (i == c) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; k++) v -= L[i * NB + k] * y[k];
            y[i] = dead[i] ? 0.0 : v / L[i * NB + i];
        }
#pragma unroll
        for (int i = NB - 1; i >= 0; i--) {
            double v = y[i];
#pragma unroll
            for (int k = i + 1; k < NB; k++) v -= L[k * NB + i] * x[k];
            x[i] = dead[i] ? 0.0 : v / L[i * NB + i];
        }
#pragma unroll
        for (int i = 0; i < NB; i++) m[i * NB + c] = x[i];
    }
    __syncthreads();
}

// The same block factorisation for larger blocks (NB > 12, whose rows do not fit a lane's
// registers), one workgroup of FACTOR_NT threads per instance over LDS copies of the blocks.
constexpr int FACTOR_NT = 256;
template <int NB>
__global__ __launch_bounds__(FACTOR_NT) void k_coarse_factor_lds(const int* __restrict__ inst_cblk,
                                                      double* __restrict__ cdiag, const double* __restrict__ clow,
                                                      double* __restrict__ cg) {
    __shared__ double Dprev[CNB_MAX * CNB_MAX], M[CNB_MAX * CNB_MAX], T[CNB_MAX * CNB_MAX],
        L[CNB_MAX * CNB_MAX], A[CNB_MAX * CNB_MAX];
    __shared__ int dead[CNB_MAX];
    const int i = blockIdx.x;
    const int k0 = inst_cblk[i], k1 = inst_cblk[i + 1];
    constexpr int nn = NB * NB, PER = (nn + FACTOR_NT - 1) / FACTOR_NT;
    // the chain's global loads of block k + 1 are issued before block k's arithmetic, so
    // no step waits on memory
    double nD[PER], nA[PER];
    auto fetch = [&](int k) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = (int)threadIdx.x + FACTOR_NT * q;
            nD[q] = e < nn ? cdiag[(int64_t)k * nn + e] : 0.0;
            nA[q] = e < nn && k > k0 ? clow[(int64_t)k * nn + e] : 0.0;  // A_(k, k-1)
        }
    };
    if (k0 < k1) fetch(k0);
    for (int k = k0; k < k1; k++) {
        double* D = cdiag + (int64_t)k * nn;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = (int)threadIdx.x + FACTOR_NT * q;
            if (e < nn) {
                M[e] = nD[q];
                A[e] = nA[q];
            }
        }
        __syncthreads();
        if (k + 1 < k1) fetch(k + 1);
        if (k > k0) {
            // G = A Dprev^-1 (Dprev holds the inverse), then M -= G A^T
            for (int e = threadIdx.x; e < nn; e += FACTOR_NT) {
                const int r = e / NB, c = e % NB;
                double v = 0.0;
                for (int t = 0; t < NB; t++) v += A[r * NB + t] * Dprev[t * NB + c];
                T[e] = v;
                cg[(int64_t)k * nn + e] = v;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < nn; e += FACTOR_NT) {
                const int r = e / NB, c = e % NB;
                double v = 0.0;
                for (int t = 0; t < NB; t++) v += T[r * NB + t] * A[c * NB + t];
                M[e] -= v;
            }
            __syncthreads();
        } else {
            for (int e = threadIdx.x; e < nn; e += FACTOR_NT) cg[(int64_t)k * nn + e] = 0.0;
        }
        // symmetrise (the probed entries agree to rounding) and invert
        for (int e = threadIdx.x; e < nn; e += FACTOR_NT) {
            const int r = e / NB, c = e % NB;
            if (c > r) {
                const double v = 0.5 * (M[r * NB + c] + M[c * NB + r]);
                T[e] = v;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nn; e += FACTOR_NT) {
            const int r = e / NB, c = e % NB;
            if (c > r) M[e] = M[c * NB + r] = T[e];
        }
        __syncthreads();
        spd_inverse<NB>(M, L, dead);
        for (int e = threadIdx.x; e < nn; e += FACTOR_NT) {
            D[e] = M[e];
            Dprev[e] = M[e];
        }
        __syncthreads();
    }
}

// coarse right-hand side: P^T r per plane (instances already converged skipped)
template <int NT>
__global__ __launch_bounds__(NT) void k_coarse_restrict(const int* __restrict__ plane_off,
                                                        const int* __restrict__ plane_nodes,
                                                        const int* __restrict__ plane_inst,
                                                        const int* __restrict__ plane_axis, int cm,
                                                        const double* __restrict__ node_r,
                                                        const int* __restrict__ I, const double* __restrict__ r,
                                                        double* __restrict__ rhs) {
    const int p = blockIdx.x;
    if (I[(int64_t)plane_inst[p] * NI + I_DONE]) return;
    double out[CM_MAX];
    plane_restrict<NT>(p, plane_off, plane_nodes, node_r, r, cm, plane_axis[p], out);
    if (threadIdx.x < cm) {
        double o = out[0];
#pragma unroll
        for (int i = 1; i < CM_MAX; i++)
            if (threadIdx.x == i) o = out[i];
        rhs[cm * (int64_t)p + threadIdx.x] = o;
    }
}

// block forward / backward sweep per instance: x = A_c^-1 rhs.  The sweep is a chain of
// small dependent matrix-vector products; wave 0 runs it out of LDS (the block size NB a
// compile-time constant, so each product's loads and broadcasts issue together) while
// waves 1-3 stage the next chunk of factor blocks and right-hand sides (double-buffered),
// so the chain never waits on a global load.  CS_WORDS doubles per buffer.
constexpr int CS_WORDS = 4608;
template <int NB, bool TRANS>
__device__ __forceinline__ double lds_matvec(const double* __restrict__ Mx, double v) {
    const int l = threadIdx.x < NB ? threadIdx.x : 0;
    double m[NB], vt[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) {
        m[t] = TRANS ? Mx[t * NB + l] : Mx[l * NB + t];
        vt[t] = __shfl(v, t, 64);
    }
    double y = 0.0;
#pragma unroll
    for (int t = 0; t < NB; t++) y += m[t] * vt[t];
    return threadIdx.x < NB ? y : 0.0;
}

// The sweep over blocks k0..k1-1 (one workgroup of 256); rhs and x indexed from block k0
// (x[(k - k0) NB + l]); rhs null: the unit vector of coarse DOF `unit` (a column of the
// inverse).
template <int NB>
__device__ __forceinline__ void coarse_sweep(int k0, int k1, const double* __restrict__ dinv,
                                             const double* __restrict__ clow, const double* __restrict__ cg,
                                             const double* __restrict__ rhs, int unit, double* __restrict__ x,
                                             double* lds) {
    constexpr int nn = NB * NB;
    const int nb = k1 - k0;
    const int tid = threadIdx.x, l = tid & 63;
    const bool sweeper = tid < 64;
    // ---- forward: z_k = b_k - G_k z_(k-1); a buffer holds CF blocks of G and then of b ----
    constexpr int CF = CS_WORDS / (nn + NB);
    const int nchf = (nb + CF - 1) / CF;
    auto stage_f = [&](int c, int buf, int t0, int nt) {
        const int kb = k0 + c * CF, ke = min(kb + CF, k1);
        double* dst = lds + buf * CS_WORDS;
        const double* src = cg + (int64_t)kb * nn;
        for (int e = t0; e < (ke - kb) * nn; e += nt) dst[e] = src[e];
        const int64_t r0 = (int64_t)(kb - k0) * NB;
        for (int e = t0; e < (ke - kb) * NB; e += nt)
            dst[CF * nn + e] = rhs ? rhs[r0 + e] : (r0 + e == unit ? 1.0 : 0.0);
    };
    stage_f(0, 0, tid, 256);
    __syncthreads();
    double z = 0.0;
    for (int c = 0; c < nchf; c++) {
        if (sweeper) {
            const double* G = lds + (c & 1) * CS_WORDS;
            const int kb = k0 + c * CF, ke = min(kb + CF, k1);
            for (int k = kb; k < ke; k++) {
                const double bk = l < NB ? G[CF * nn + (k - kb) * NB + l] : 0.0;
                const double gz = k > k0 ? lds_matvec<NB, false>(G + (k - kb) * nn, z) : 0.0;
                z = bk - gz;
                if (l < NB) x[(int64_t)(k - k0) * NB + l] = z;
            }
        } else if (c + 1 < nchf) {
            stage_f(c + 1, (c + 1) & 1, tid - 64, 192);
        }
        __syncthreads();
    }
    // ---- backward: x_k = D_k^-1 (z_k - A_(k+1,k)^T x_(k+1)); a buffer holds CB blocks of
    // (D_k^-1, A_(k+1,k)) and then their z_k, from the last block down ----
    constexpr int CB = CS_WORDS / (2 * nn + NB);
    const int nchb = (nb + CB - 1) / CB;
    auto stage_b = [&](int c, int buf, int t0, int nt) {
        const int ke = k1 - c * CB, kb = max(ke - CB, k0);
        double* dst = lds + buf * CS_WORDS;
        for (int e = t0; e < (ke - kb) * nn; e += nt) {
            const int k = kb + e / nn, o = e - (k - kb) * nn;
            dst[2 * (k - kb) * nn + o] = dinv[(int64_t)k * nn + o];
            dst[(2 * (k - kb) + 1) * nn + o] = k + 1 < k1 ? clow[(int64_t)(k + 1) * nn + o] : 0.0;
        }
        const double* zs = x + (int64_t)(kb - k0) * NB;
        for (int e = t0; e < (ke - kb) * NB; e += nt) dst[2 * CB * nn + e] = zs[e];
    };
    stage_b(0, 0, tid, 256);
    __syncthreads();
    double xn = 0.0;
    for (int c = 0; c < nchb; c++) {
        if (sweeper) {
            const double* B = lds + (c & 1) * CS_WORDS;
            const int ke = k1 - c * CB, kb = max(ke - CB, k0);
            for (int k = ke - 1; k >= kb; k--) {
                double t = l < NB ? B[2 * CB * nn + (k - kb) * NB + l] : 0.0;
                if (k < k1 - 1) t -= lds_matvec<NB, true>(B + (2 * (k - kb) + 1) * nn, xn);
                xn = lds_matvec<NB, false>(B + 2 * (k - kb) * nn, t);
                if (l < NB) x[(int64_t)(k - k0) * NB + l] = xn;
            }
        } else if (c + 1 < nchb) {
            stage_b(c + 1, (c + 1) & 1, tid - 64, 192);
        }
        __syncthreads();
    }
}

// x = A_c^-1 rhs per instance by the sweep; instances with a dense inverse (dense_off >= 0)
// are left to k_coarse_dense_apply
template <int NB>
__global__ __launch_bounds__(256) void k_coarse_solve(const int* __restrict__ inst_cblk, const int* __restrict__ I,
                                                      const int64_t* __restrict__ dense_off,
                                                      const double* __restrict__ dinv, const double* __restrict__ clow,
                                                      const double* __restrict__ cg, const double* __restrict__ rhs,
                                                      double* __restrict__ x) {
    extern __shared__ double lds[];  // [2][CS_WORDS]
    const int i = blockIdx.x;
    if (I[(int64_t)i * NI + I_DONE] || (dense_off && dense_off[i] >= 0)) return;
    const int k0 = inst_cblk[i], k1 = inst_cblk[i + 1];
    coarse_sweep<NB>(k0, k1, dinv, clow, cg, rhs + (int64_t)k0 * NB, -1, x + (int64_t)k0 * NB, lds);
}

// column blockIdx.x of the instance's A_c^-1 (n_c = (k1 - k0) NB rows) into Ainv
template <int NB>
__global__ __launch_bounds__(256) void k_coarse_inverse(int k0, int k1, const double* __restrict__ dinv,
                                                        const double* __restrict__ clow, const double* __restrict__ cg,
                                                        double* __restrict__ Ainv) {
    extern __shared__ double lds[];  // [2][CS_WORDS]
    const int64_t nc = (int64_t)(k1 - k0) * NB;
    coarse_sweep<NB>(k0, k1, dinv, clow, cg, nullptr, (int)blockIdx.x, Ainv + nc * blockIdx.x, lds);
}

// x = A_c^-1 rhs for the instances with a dense inverse: one wave per coarse row, four rows
// per workgroup (the inverse is symmetric, so row r is read as its contiguous column r)
__global__ __launch_bounds__(256) void k_coarse_dense_apply(int NB, int n_rows, const int* __restrict__ rows,
                                                            const int* __restrict__ row_inst,
                                                            const int* __restrict__ inst_cblk,
                                                            const int64_t* __restrict__ dense_off,
                                                            const int* __restrict__ I, const double* __restrict__ Ainv,
                                                            const double* __restrict__ rhs, double* __restrict__ x) {
    const int w = (int)(blockIdx.x * 4 + (threadIdx.x >> 6)), l = threadIdx.x & 63;
    if (w >= n_rows) return;
    const int g = rows[w], i = row_inst[w];
    if (I[(int64_t)i * NI + I_DONE]) return;
    const int64_t base = (int64_t)inst_cblk[i] * NB, nc = (int64_t)(inst_cblk[i + 1] - inst_cblk[i]) * NB;
    const double* A = Ainv + dense_off[i] + (g - base) * nc;
    const double* b = rhs + base;
    // four independent sums in flight, added in a fixed order
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t j = l;
    for (; j + 192 < nc; j += 256) {
        a0 += A[j] * b[j];
        a1 += A[j + 64] * b[j + 64];
        a2 += A[j + 128] * b[j + 128];
        a3 += A[j + 192] * b[j + 192];
    }
    for (; j < nc; j += 64) a0 += A[j] * b[j];
    double acc = (a0 + a1) + (a2 + a3);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (l == 0) x[g] = acc;
}

int launch_coarse_solve(int NB, int n_inst, const int* inst_cblk, const int* I, const int64_t* dense_off,
                        const double* dinv, const double* clow, const double* cg, const double* rhs, double* x,
                        hipStream_t s) {
    const size_t lds = 2 * CS_WORDS * sizeof(double);
    if (NB == 6)
        k_coarse_solve<6><<<n_inst, 256, lds, s>>>(inst_cblk, I, dense_off, dinv, clow, cg, rhs, x);
    else if (NB == 11)
        k_coarse_solve<11><<<n_inst, 256, lds, s>>>(inst_cblk, I, dense_off, dinv, clow, cg, rhs, x);
    else if (NB == 12)
        k_coarse_solve<12><<<n_inst, 256, lds, s>>>(inst_cblk, I, dense_off, dinv, clow, cg, rhs, x);
    else if (NB == 22)
        k_coarse_solve<22><<<n_inst, 256, lds, s>>>(inst_cblk, I, dense_off, dinv, clow, cg, rhs, x);
    else
        k_coarse_solve<18><<<n_inst, 256, lds, s>>>(inst_cblk, I, dense_off, dinv, clow, cg, rhs, x);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

int launch_coarse_inverse(int NB, int k0, int k1, const double* dinv, const double* clow, const double* cg,
                          double* Ainv, hipStream_t s) {
    const size_t lds = 2 * CS_WORDS * sizeof(double);
    const unsigned cols = (unsigned)((k1 - k0) * NB);
    if (NB == 6)
        k_coarse_inverse<6><<<cols, 256, lds, s>>>(k0, k1, dinv, clow, cg, Ainv);
    else if (NB == 11)
        k_coarse_inverse<11><<<cols, 256, lds, s>>>(k0, k1, dinv, clow, cg, Ainv);
    else if (NB == 12)
        k_coarse_inverse<12><<<cols, 256, lds, s>>>(k0, k1, dinv, clow, cg, Ainv);
    else if (NB == 22)
        k_coarse_inverse<22><<<cols, 256, lds, s>>>(k0, k1, dinv, clow, cg, Ainv);
    else
        k_coarse_inverse<18><<<cols, 256, lds, s>>>(k0, k1, dinv, clow, cg, Ainv);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

}  // namespace tlfea_hip

using namespace tlfea_hip;

namespace {

template <typename T>
int dalloc(T** p, size_t n) {
    if (n == 0) n = 1;
    if (hipMalloc((void**)p, sizeof(T) * n) != hipSuccess) {
        set_error("device allocation of %zu bytes failed", sizeof(T) * n);
        return TLFEA_EHIP;
    }
    return TLFEA_OK;
}

void free_batch(tlfea_batch* b) {
    if (!b) return;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->plan) tlfea_force_plan_destroy(b->plan);
    void* ptrs[] = {b->d_conn, b->d_X, b->d_Y, b->d_Z, b->d_mask, b->d_elem_inst, b->d_jinv, b->d_detJ,
                    b->d_jinv32, b->d_detJ32, b->d_lam, b->d_mu, b->d_lam32, b->d_mu32, b->d_chunk_inst,
                    b->d_chunk_beg, b->d_chunk_end, b->d_inst_chunk_off, b->d_r, b->d_p, b->d_q, b->d_z,
                    b->d_r32, b->d_z32, b->d_d32, b->d_q32, b->d_part, b->d_s, b->d_i, b->d_all, b->d_cheb,
                    b->d_rold, b->d_dinv32, b->d_blk_list, b->d_chunk_list, b->d_inst_list, b->d_mc, b->d_mc32,
                    b->d_ux, b->d_uy, b->d_uz, b->d_node_plane, b->d_node_r, b->d_plane_off,
                    b->d_plane_nodes, b->d_plane_inst, b->d_plane_local, b->d_inst_cblk, b->d_cdiag,
                    b->d_clow, b->d_cg, b->d_crhs, b->d_cx, b->d_nvk, b->d_node_inst, b->d_npart, b->d_ns, b->d_nmc, b->d_nrs, b->d_cdense_off, b->d_cdense,
                    b->d_cd_row, b->d_cd_inst, b->d_cinst_pn, b->d_usol, b->d_plane_axis};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->s_own) (void)hipStreamDestroy(b->s_own);
    delete b;
}

Chunks chunks(const tlfea_batch* b) {
    return Chunks{b->d_chunk_inst, b->d_chunk_beg, b->d_chunk_end, b->compact ? b->d_chunk_list : nullptr};
}
int n_chunk_launch(const tlfea_batch* b) { return b->compact ? b->n_chunk_list : b->n_chunks; }
Insts insts(const tlfea_batch* b) {
    return b->compact ? Insts{b->n_inst_list, b->d_inst_list} : Insts{b->n_inst, nullptr};
}

void drop_graph(tlfea_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
}

// Restricts the launches to the instances not yet done (host flags `done`).
int compact_lists(tlfea_batch* b, const std::vector<int>& done) {
    std::vector<int> il, cl, bl;
    int64_t elems = 0;
    for (int i = 0; i < b->n_inst; i++) {
        if (done[i]) continue;
        il.push_back(i);
        for (int c = b->h_inst_chunk_off[i]; c < b->h_inst_chunk_off[i + 1]; c++) cl.push_back(c);
        const int b0 = b->h_elem_off[i] / PLAN_BLOCK, b1 = (b->h_elem_off[i + 1] - 1) / PLAN_BLOCK;
        for (int k = b0; k <= b1; k++)
            if (bl.empty() || bl.back() < k) bl.push_back(k);
        elems += b->h_elem_off[i + 1] - b->h_elem_off[i];
    }
    if (il.empty()) return TLFEA_OK;
    TLFEA_HIP_TRY(hipMemcpy(b->d_inst_list, il.data(), sizeof(int) * il.size(), hipMemcpyHostToDevice));
    TLFEA_HIP_TRY(hipMemcpy(b->d_chunk_list, cl.data(), sizeof(int) * cl.size(), hipMemcpyHostToDevice));
    TLFEA_HIP_TRY(hipMemcpy(b->d_blk_list, bl.data(), sizeof(int) * bl.size(), hipMemcpyHostToDevice));
    b->n_inst_list = (int)il.size();
    b->n_chunk_list = (int)cl.size();
    b->n_blk_list = (int)bl.size();
    b->compact = true;
    b->launched_elems = elems;
    drop_graph(b);
    return TLFEA_OK;
}

void full_lists(tlfea_batch* b) {
    if (b->compact) drop_graph(b);
    b->compact = false;
    b->launched_elems = b->n_elem;
}

// (Re)builds the chunk tables for a chunk size.
int build_chunks(tlfea_batch* b, int chunk_nodes) {
    std::vector<int> ci, cb, ce, ico(b->n_inst + 1, 0);
    for (int i = 0; i < b->n_inst; i++) {
        ico[i] = (int)ci.size();
        for (int n0 = b->h_node_off[i]; n0 < b->h_node_off[i + 1]; n0 += chunk_nodes) {
            ci.push_back(i);
            cb.push_back(n0);
            ce.push_back(std::min(n0 + chunk_nodes, b->h_node_off[i + 1]));
        }
    }
    ico[b->n_inst] = (int)ci.size();
    if (ci.empty()) {  // keep one empty chunk so launches are never empty
        ci.push_back(0);
        cb.push_back(0);
        ce.push_back(0);
    }
    for (int* p : {b->d_chunk_inst, b->d_chunk_beg, b->d_chunk_end, b->d_inst_chunk_off})
        if (p) (void)hipFree(p);
    if (b->d_part) (void)hipFree(b->d_part);
    b->d_chunk_inst = b->d_chunk_beg = b->d_chunk_end = b->d_inst_chunk_off = nullptr;
    b->d_part = nullptr;
    int rc;
    if ((rc = dalloc(&b->d_chunk_inst, ci.size())) || (rc = dalloc(&b->d_chunk_beg, cb.size())) ||
        (rc = dalloc(&b->d_chunk_end, ce.size())) || (rc = dalloc(&b->d_inst_chunk_off, ico.size())) ||
        (rc = dalloc(&b->d_part, 2 * ci.size())))
        return rc;
    TLFEA_HIP_TRY(hipMemcpy(b->d_chunk_inst, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice));
    TLFEA_HIP_TRY(hipMemcpy(b->d_chunk_beg, cb.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice));
    TLFEA_HIP_TRY(hipMemcpy(b->d_chunk_end, ce.data(), sizeof(int) * ce.size(), hipMemcpyHostToDevice));
    TLFEA_HIP_TRY(hipMemcpy(b->d_inst_chunk_off, ico.data(), sizeof(int) * ico.size(), hipMemcpyHostToDevice));
    if (b->d_chunk_list) (void)hipFree(b->d_chunk_list);
    b->d_chunk_list = nullptr;
    if ((rc = dalloc(&b->d_chunk_list, ci.size()))) return rc;
    b->h_inst_chunk_off = ico;
    b->compact = false;
    b->chunk_nodes = chunk_nodes;
    b->n_chunks = (int)ci.size();
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return TLFEA_OK;
}

// FP64 batched operator q = (k_scale K + c M) p (masked; the mass term when set up),
// chunk partials of p.q into d_part.
// the kernels' (x, X0) pair: (u, null) -- the displacement itself -- once moved, else the
// rest positions twice
const double* pos_x(const tlfea_batch* b) { return b->moved ? b->d_ux : b->d_X; }
const double* pos_y(const tlfea_batch* b) { return b->moved ? b->d_uy : b->d_Y; }
const double* pos_z(const tlfea_batch* b) { return b->moved ? b->d_uz : b->d_Z; }
const double* rest_x(const tlfea_batch* b) { return b->moved ? nullptr : b->d_X; }
const double* rest_y(const tlfea_batch* b) { return b->moved ? nullptr : b->d_Y; }
const double* rest_z(const tlfea_batch* b) { return b->moved ? nullptr : b->d_Z; }

int apply64(tlfea_batch* b, const double* p, double* q, const int* iflags, hipStream_t s, double k_scale = 1.0,
            bool mass = true, bool masked = true) {
    TanMat<double> tm{0.0, 0.0, b->d_elem_inst, b->d_lam, b->d_mu, iflags, NI};
    tm.inst_mc = (mass && b->mass_on) ? b->d_mc : nullptr;
    tm.mass_exact = b->mass_exact;
    tm.k_scale = k_scale;
    tm.at_rest = b->moved ? 0 : 1;  // the tangent at the rest positions unless moved
    tm.affine = b->affine;
    const unsigned char* mask = masked ? b->d_mask : nullptr;
    if (int rc = launch_tangent_partials<double>(b->plan, pos_x(b), pos_y(b), pos_z(b), rest_x(b), rest_y(b),
                                                 rest_z(b), b->d_jinv,
                                                 b->d_detJ, b->quad, tm, mask, p, nullptr, s,
                                                 b->compact ? b->d_blk_list : nullptr, b->n_blk_list))
        return rc;
    k_bgather<double><<<n_chunk_launch(b), BT, 0, s>>>(chunks(b), b->plan->d_g_off, b->plan->d_g_idx,
                                                       b->plan->d_part, mask, p, q, b->d_part, iflags);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// Block partials of the FP32 batched operator (K + c M) p (masked; the mass term when set up).
int tangent32_partials(tlfea_batch* b, const float* p, const int* iflags, hipStream_t s) {
    TanMat<float> tm{0.0f, 0.0f, b->d_elem_inst, b->d_lam32, b->d_mu32, iflags, NI};
    tm.inst_mc = b->mass_on ? b->d_mc32 : nullptr;
    tm.mass_exact = b->mass_exact;
    tm.at_rest = b->moved ? 0 : 1;
    tm.affine = b->affine;
    return launch_tangent_partials<float>(b->plan, pos_x(b), pos_y(b), pos_z(b), rest_x(b), rest_y(b), rest_z(b),
                                          b->d_jinv32, b->d_detJ32, b->quad, tm, b->d_mask, p, nullptr, s,
                                          b->compact ? b->d_blk_list : nullptr, b->n_blk_list);
}

// FP32 batched operator q = (K + c M) p (masked; the mass term when set up), no dot partials.
int apply32(tlfea_batch* b, const float* p, float* q, const int* iflags, hipStream_t s) {
    if (int rc = tangent32_partials(b, p, iflags, s)) return rc;
    k_bgather<float><<<n_chunk_launch(b), BT, 0, s>>>(chunks(b), b->plan->d_g_off, b->plan->d_g_idx,
                                                reinterpret_cast<const float*>(b->plan->d_part), b->d_mask, p, q,
                                                nullptr, iflags);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// z = Chebyshev(r) for the active instances, plus the coarse correction P A_c^-1 P^T r
// (FP64) when it is set up -- the additive two-level preconditioner -- then z.r / z.r_old
// chunk partials.
// rhs_c = P^T r per plane: four waves per plane on large planes, one on small ones
void launch_restrict(tlfea_batch* b, const int* plane_off, const int* plane_nodes, const int* plane_inst,
                     const double* node_r, const int* I, const double* r, double* rhs, hipStream_t s) {
    if (b->restrict_wide)
        k_coarse_restrict<256><<<b->n_planes, 256, 0, s>>>(plane_off, plane_nodes, plane_inst, b->d_plane_axis,
                                                           b->cm, node_r, I, r, rhs);
    else
        k_coarse_restrict<64><<<b->n_planes, 64, 0, s>>>(plane_off, plane_nodes, plane_inst, b->d_plane_axis, b->cm,
                                                         node_r, I, r, rhs);
}

// x_c = A_c^-1 rhs_c for every active instance: the block sweep, or the dense inverse
int coarse_solve_all(tlfea_batch* b, hipStream_t s) {
    if (b->n_cd_inst < b->n_inst)  // some instances sweep
        if (int rc = launch_coarse_solve(b->cnb, b->n_inst, b->d_inst_cblk, b->d_i, b->d_cdense_off, b->d_cdiag,
                                         b->d_clow, b->d_cg, b->d_crhs, b->d_cx, s))
            return rc;
    if (b->n_cd_rows > 0) {
        k_coarse_dense_apply<<<(b->n_cd_rows + 3) / 4, 256, 0, s>>>(b->cnb, b->n_cd_rows, b->d_cd_row, b->d_cd_inst,
                                                                    b->d_inst_cblk, b->d_cdense_off, b->d_i,
                                                                    b->d_cdense, b->d_crhs, b->d_cx);
        TLFEA_LAUNCH_CHECK();
    }
    return TLFEA_OK;
}

int precondition(tlfea_batch* b, hipStream_t s) {
    const Chunks ch = chunks(b);
    CoarseProlong cp{nullptr, nullptr, nullptr};
    if (b->coarse_on) {
        launch_restrict(b, b->d_plane_off, b->d_plane_nodes, b->d_plane_inst, b->d_node_r, b->d_i, b->d_r,
                        b->d_crhs, s);
        if (int rc = coarse_solve_all(b, s)) return rc;
        cp = CoarseProlong{b->d_node_plane, b->d_node_r, b->d_cx, b->cm, b->d_plane_axis};
    }
    k_cheb_init<<<n_chunk_launch(b), BT, 0, s>>>(ch, b->d_r, b->d_dinv32, b->d_cheb, b->d_i, b->d_r32, b->d_d32,
                                                b->d_z32);
    for (int k = 1; k < b->cheb_degree; k++) {
        if (int rc = tangent32_partials(b, b->d_d32, b->d_i, s)) return rc;
        k_bgather_cheb<<<n_chunk_launch(b), BT, 0, s>>>(ch, b->plan->d_g_off, b->plan->d_g_idx,
                                                       reinterpret_cast<const float*>(b->plan->d_part), b->d_mask, k,
                                                       b->d_dinv32, b->d_cheb, b->d_i, b->d_r32, b->d_d32, b->d_z32);
    }
    k_cheb_fin<<<n_chunk_launch(b), BT, 0, s>>>(ch, b->d_z32, b->d_r, b->d_rold, b->d_i, b->d_z, b->d_part, cp);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// One flexible preconditioned CG iteration for every active instance.
int pcg_iteration(tlfea_batch* b, double* u, double tol, hipStream_t s) {
    const Chunks ch = chunks(b);
    if (int rc = apply64(b, b->d_p, b->d_q, b->d_i, s)) return rc;
    const Insts in = insts(b);
    k_breduce_alpha<<<in.n, 64, 0, s>>>(in, b->d_inst_chunk_off, b->d_part, b->d_s, b->d_i);
    k_bupdate_keep<<<n_chunk_launch(b), BT, 0, s>>>(ch, u, b->d_r, b->d_rold, b->d_p, b->d_q, b->d_s, b->d_i, b->d_part,
                                                   b->d_all);
    k_breduce_check<<<in.n, 64, 0, s>>>(in, b->d_inst_chunk_off, b->d_part, b->d_s, b->d_i, b->d_all, tol);
    if (int rc = precondition(b, s)) return rc;
    k_breduce_fbeta<<<in.n, 64, 0, s>>>(in, b->d_inst_chunk_off, b->d_part, b->d_s, b->d_i);
    k_bdir<<<n_chunk_launch(b), BT, 0, s>>>(ch, b->d_p, b->d_z, b->d_s, b->d_i);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// One single-reduction CG iteration for every active instance (3 launches, no host
// synchronisation); par = parity of the iteration index.  d_q holds w = K r, d_z s = K p.
int cg_iteration(tlfea_batch* b, double* u, double tol, hipStream_t s, int par) {
    const Chunks ch = chunks(b);
    TanMat<double> tm{0.0, 0.0, b->d_elem_inst, b->d_lam, b->d_mu, b->d_i, NI};
    tm.inst_mc = b->mass_on ? b->d_mc : nullptr;
    tm.mass_exact = b->mass_exact;
    tm.at_rest = b->moved ? 0 : 1;
    tm.affine = b->affine;
    if (int rc = launch_tangent_partials<double>(b->plan, pos_x(b), pos_y(b), pos_z(b), rest_x(b), rest_y(b),
                                                 rest_z(b), b->d_jinv, b->d_detJ, b->quad, tm, b->d_mask, b->d_r,
                                                 nullptr, s, b->compact ? b->d_blk_list : nullptr, b->n_blk_list))
        return rc;
    k_bgather_rw<<<n_chunk_launch(b), BT, 0, s>>>(ch, b->plan->d_g_off, b->plan->d_g_idx, b->plan->d_part, b->d_mask,
                                                 b->d_r, b->d_q, b->d_part, b->d_i, b->d_all);
    k_bcg_update<<<n_chunk_launch(b), BT, 0, s>>>(ch, b->d_inst_chunk_off, b->d_part, u, b->d_r, b->d_p, b->d_z,
                                                 b->d_q, b->d_s, b->d_i, b->d_all, tol, par);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

}  // namespace

extern "C" int tlfea_batch_create(int nn, int n_inst, const int* elem_off, const int* node_off, const int* conn,
                                  const double* x, const double* y, const double* z, const double* lambda,
                                  const double* mu, const unsigned char* fixed_dof_mask, tlfea_batch** out,
                                  void* stream) {
    TLFEA_REQUIRE(out != nullptr, "null output handle");
    *out = nullptr;
    TLFEA_REQUIRE(nn == 10 || nn == 4, "nn must be 10 or 4");
    TLFEA_REQUIRE(n_inst > 0 && elem_off && node_off && conn && x && y && z && lambda && mu && fixed_dof_mask,
                  "bad argument");
    TLFEA_REQUIRE(elem_off[0] == 0 && node_off[0] == 0, "offsets must start at 0");
    for (int i = 0; i < n_inst; i++)
        TLFEA_REQUIRE(elem_off[i + 1] > elem_off[i] && node_off[i + 1] > node_off[i],
                      "every instance needs at least one element and one node");
    const int n_elem = elem_off[n_inst], n_nodes = node_off[n_inst];
    TLFEA_REQUIRE(n_elem > 0 && n_nodes > 0, "empty batch");
    hipStream_t s = as_stream(stream);
    auto* b = new tlfea_batch();
    b->nn = nn;
    b->n_inst = n_inst;
    b->n_elem = n_elem;
    b->n_nodes = n_nodes;
    b->h_elem_off.assign(elem_off, elem_off + n_inst + 1);
    b->h_node_off.assign(node_off, node_off + n_inst + 1);
    int rc = tlfea_quadrature(nn, &b->quad);
    const int nq = b->quad.n;
    const int64_t pe = padded(n_elem);
    const size_t nd = 3 * (size_t)n_nodes;
    auto fail = [&](int code) {
        free_batch(b);
        return code;
    };
    if (rc) return fail(rc);
    if ((rc = dalloc(&b->d_conn, (size_t)n_elem * nn)) || (rc = dalloc(&b->d_X, n_nodes)) ||
        (rc = dalloc(&b->d_Y, n_nodes)) || (rc = dalloc(&b->d_Z, n_nodes)) || (rc = dalloc(&b->d_mask, nd)) ||
        (rc = dalloc(&b->d_elem_inst, n_elem)) || (rc = dalloc(&b->d_jinv, (size_t)pe * nq * 9)) ||
        (rc = dalloc(&b->d_detJ, (size_t)pe * nq)) || (rc = dalloc(&b->d_jinv32, (size_t)pe * nq * 9)) ||
        (rc = dalloc(&b->d_detJ32, (size_t)pe * nq)) || (rc = dalloc(&b->d_lam, n_inst)) ||
        (rc = dalloc(&b->d_mu, n_inst)) || (rc = dalloc(&b->d_lam32, n_inst)) || (rc = dalloc(&b->d_mu32, n_inst)) ||
        (rc = dalloc(&b->d_r, nd)) || (rc = dalloc(&b->d_p, nd)) || (rc = dalloc(&b->d_q, nd)) ||
        (rc = dalloc(&b->d_z, nd)) || (rc = dalloc(&b->d_r32, nd)) || (rc = dalloc(&b->d_z32, nd)) ||
        (rc = dalloc(&b->d_d32, nd)) || (rc = dalloc(&b->d_q32, nd)) || (rc = dalloc(&b->d_s, (size_t)n_inst * NS)) ||
        (rc = dalloc(&b->d_i, (size_t)n_inst * NI)) || (rc = dalloc(&b->d_all, 4)) ||
        (rc = dalloc(&b->d_cheb, (size_t)n_inst * NC)) || (rc = dalloc(&b->d_rold, nd)) ||
        (rc = dalloc(&b->d_dinv32, nd)) || (rc = dalloc(&b->d_usol, nd)))
        return fail(rc);
    if (hipMemcpyAsync(b->d_conn, conn, sizeof(int) * (size_t)n_elem * nn, hipMemcpyDeviceToDevice, s) ||
        hipMemcpyAsync(b->d_X, x, sizeof(double) * n_nodes, hipMemcpyDeviceToDevice, s) ||
        hipMemcpyAsync(b->d_Y, y, sizeof(double) * n_nodes, hipMemcpyDeviceToDevice, s) ||
        hipMemcpyAsync(b->d_Z, z, sizeof(double) * n_nodes, hipMemcpyDeviceToDevice, s) ||
        hipMemcpyAsync(b->d_mask, fixed_dof_mask, nd, hipMemcpyDeviceToDevice, s) ||
        hipMemcpyAsync(b->d_lam, lambda, sizeof(double) * n_inst, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(b->d_mu, mu, sizeof(double) * n_inst, hipMemcpyHostToDevice, s)) {
        set_error("batch input copy failed");
        return fail(TLFEA_EHIP);
    }
    // element -> instance map (offsets staged through the r workspace)
    int* d_eoff = reinterpret_cast<int*>(b->d_r);
    if (hipMemcpyAsync(d_eoff, elem_off, sizeof(int) * (n_inst + 1), hipMemcpyHostToDevice, s)) {
        set_error("batch offset copy failed");
        return fail(TLFEA_EHIP);
    }
    k_elem_inst<<<n_inst, 256, 0, s>>>(n_inst, d_eoff, b->d_elem_inst);
    if ((rc = launch_precompute(nn, n_elem, b->d_conn, b->d_X, b->d_Y, b->d_Z, b->quad, nullptr, b->d_detJ, b->d_jinv,
                                s)))
        return fail(rc);
    if ((rc = tlfea_jinv_affine(nn, n_elem, b->d_jinv, b->d_detJ, &b->affine, s))) return fail(rc);
    k_d2f<<<grid_for((int64_t)pe * nq * 9, 256), 256, 0, s>>>((int64_t)pe * nq * 9, b->d_jinv, b->d_jinv32);
    k_d2f<<<grid_for((int64_t)pe * nq, 256), 256, 0, s>>>((int64_t)pe * nq, b->d_detJ, b->d_detJ32);
    k_d2f<<<grid_for(n_inst, 256), 256, 0, s>>>(n_inst, b->d_lam, b->d_lam32);
    k_d2f<<<grid_for(n_inst, 256), 256, 0, s>>>(n_inst, b->d_mu, b->d_mu32);
    if (hipGetLastError() != hipSuccess) {
        set_error("batch setup kernel launch failed");
        return fail(TLFEA_EHIP);
    }
    if ((rc = tlfea_force_plan_create(nn, n_elem, n_nodes, b->d_conn, &b->plan, stream))) return fail(rc);
    if ((rc = dalloc(&b->d_blk_list, b->plan->n_blocks)) || (rc = dalloc(&b->d_inst_list, n_inst))) return fail(rc);
    b->launched_elems = n_elem;
    if ((rc = build_chunks(b, 512))) return fail(rc);
    if (hipStreamCreateWithFlags(&b->s_own, hipStreamNonBlocking) != hipSuccess) {
        set_error("stream creation failed");
        return fail(TLFEA_EHIP);
    }
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    *out = b;
    return TLFEA_OK;
}

extern "C" int tlfea_batch_destroy(tlfea_batch* b) {
    free_batch(b);
    return TLFEA_OK;
}

extern "C" int tlfea_batch_sizes(const tlfea_batch* b, int* n_inst, int* n_elem, int* n_nodes, int* n_chunks) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    if (n_inst) *n_inst = b->n_inst;
    if (n_elem) *n_elem = b->n_elem;
    if (n_nodes) *n_nodes = b->n_nodes;
    if (n_chunks) *n_chunks = b->n_chunks;
    return TLFEA_OK;
}

extern "C" int tlfea_batch_set_chunk(tlfea_batch* b, int chunk_nodes) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    TLFEA_REQUIRE(chunk_nodes >= 64 && chunk_nodes <= (1 << 20), "chunk size out of range [64, 2^20]");
    return build_chunks(b, chunk_nodes);
}

extern "C" int tlfea_batch_autotune(tlfea_batch* b, const double* f, double* u, int trial_iters, double tol,
                                    int full_solves, tlfea_batch_tuning* out, void* stream) {
    TLFEA_REQUIRE(b && f && u && out, "null argument");
    TLFEA_REQUIRE(trial_iters >= 8 && trial_iters <= 65536, "trial iterations out of range [8, 65536]");
    TLFEA_REQUIRE(!full_solves || tol > 0.0, "full-solve trials need a positive tolerance");
    const std::vector<int> chunk_set = {256, 512, 1024, 2048, 4096};
    const std::vector<int> poll_set = {16, 32, 64};
    const std::vector<double> ratio_set = {0.35, 0.5, 0.65, 0.8, 0.9};
    std::vector<int> it(b->n_inst), st(b->n_inst);
    std::vector<double> rr(b->n_inst);
    int total = 0;
    using clock = std::chrono::steady_clock;
    // wall time of one solve of n iterations (tol 0: no instance finishes, nothing is
    // compacted); the call returns after its final device-to-host copies
    auto timed = [&](int n, int poll, double t, double& sec) -> int {
        const auto t0 = clock::now();
        const int rc = tlfea_batch_cg_solve(b, f, u, 0, t, n, poll, 1, 0, it.data(), rr.data(), st.data(), &total,
                                            stream);
        sec = std::chrono::duration<double>(clock::now() - t0).count();
        return (rc == TLFEA_OK || rc == TLFEA_ENOCONV) ? TLFEA_OK : rc;
    };
    const double keep_ratio = b->compact_ratio;
    b->compact_ratio = 0.0;
    double best = 1e300;
    int best_chunk = b->chunk_nodes, best_poll = b->poll_default;
    for (int chunk : chunk_set) {
        if (int rc = build_chunks(b, chunk)) return rc;
        for (int poll : poll_set) {
            // per-iteration time from runs of n and 2n iterations: the host round trips at
            // the ends of a solve cancel in the difference.  Each length is timed twice and
            // the faster run kept (the first run after a chunk or poll change pays one-time
            // costs, the graph capture among them); a difference lost in the noise falls
            // back to the 2n run's mean, an upper bound.
            const int n = poll * ((trial_iters + poll - 1) / poll);
            double t1 = 1e300, t2 = 1e300;
            for (int rep = 0; rep < 2; rep++) {
                double a = 0.0, c = 0.0;
                if (int rc = timed(n, poll, 0.0, a)) return rc;
                if (int rc = timed(2 * n, poll, 0.0, c)) return rc;
                t1 = std::min(t1, a);
                t2 = std::min(t2, c);
            }
            const double per = t2 > t1 ? (t2 - t1) / n : t2 / (2.0 * n);
            if (per < best) {
                best = per;
                best_chunk = chunk;
                best_poll = poll;
            }
        }
    }
    if (int rc = build_chunks(b, best_chunk)) return rc;
    b->poll_default = best_poll;
    b->compact_ratio = keep_ratio;
    out->chunk_nodes = best_chunk;
    out->poll = best_poll;
    out->us_per_iter = best * 1e6;
    out->compaction = keep_ratio;
    out->solve_s = 0.0;
    if (full_solves) {
        // the compaction ratio decides the tail of uneven batches: full solves to tol
        double best_s = 1e300, best_r = keep_ratio;
        for (double r : ratio_set) {
            b->compact_ratio = r;
            double sec = 0.0;
            if (int rc = timed(1 << 30, best_poll, tol, sec)) return rc;
            if (sec < best_s) {
                best_s = sec;
                best_r = r;
            }
        }
        b->compact_ratio = best_r;
        out->compaction = best_r;
        out->solve_s = best_s;
    }
    return TLFEA_OK;
}

extern "C" int tlfea_batch_precond_setup(tlfea_batch* b, int degree, double ratio, int power_iters,
                                         double safety, double* lambda_max, void* stream) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    TLFEA_REQUIRE(degree >= 0 && degree <= 64, "Chebyshev degree out of range [0, 64]");
    TLFEA_REQUIRE(degree == 0 || (ratio > 1.0 && power_iters >= 1 && safety >= 1.0), "bad Chebyshev parameters");
    b->cheb_degree = 0;
    drop_graph(b);
    full_lists(b);
    if (degree == 0) return TLFEA_OK;
    hipStream_t s = as_stream(stream);
    const Chunks ch = chunks(b);
    const int64_t nd = 3 * (int64_t)b->n_nodes;
    // diagonal of K + c M at the current configuration: partials, then the gather with p = 1
    TanMat<double> tm{0.0, 0.0, b->d_elem_inst, b->d_lam, b->d_mu, nullptr, NI};
    tm.inst_mc = b->mass_on ? b->d_mc : nullptr;
    tm.mass_exact = b->mass_exact;
    if (int rc = launch_tangent_diag_partials(b->plan, pos_x(b), pos_y(b), pos_z(b), rest_x(b), rest_y(b), rest_z(b),
                                              b->d_jinv,
                                              b->d_detJ, b->quad, tm, s))
        return rc;
    k_fill<<<grid_for(nd, 256), 256, 0, s>>>(nd, 1.0, b->d_z);
    k_bgather<double><<<b->n_chunks, BT, 0, s>>>(ch, b->plan->d_g_off, b->plan->d_g_idx, b->plan->d_part, b->d_mask,
                                                 b->d_z, b->d_q, nullptr, nullptr);
    k_inv_diag<<<grid_for(nd, 256), 256, 0, s>>>(nd, b->d_q, b->d_dinv32);
    // power iteration on D^-1 K
    k_pow_start<<<b->n_chunks, BT, 0, s>>>(ch, b->d_mask, b->d_d32);
    for (int it = 0; it < power_iters; it++) {
        if (int rc = apply32(b, b->d_d32, b->d_q32, nullptr, s)) return rc;
        k_pow_scale<<<b->n_chunks, BT, 0, s>>>(ch, b->d_q32, b->d_dinv32, b->d_mask, b->d_z32, b->d_part);
        k_pow_reduce<<<b->n_inst, 64, 0, s>>>(b->n_inst, b->d_inst_chunk_off, b->d_part, b->d_cheb);
        k_pow_norm<<<b->n_chunks, BT, 0, s>>>(ch, b->d_z32, b->d_cheb, b->d_d32);
    }
    k_cheb_params<<<grid_for(b->n_inst, 256), 256, 0, s>>>(b->n_inst, safety, ratio, b->d_cheb);
    TLFEA_LAUNCH_CHECK();
    if (lambda_max) {
        std::vector<double> C((size_t)b->n_inst * NC);
        TLFEA_HIP_TRY(hipMemcpyAsync(C.data(), b->d_cheb, sizeof(double) * C.size(), hipMemcpyDeviceToHost, s));
        TLFEA_HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < b->n_inst; i++) lambda_max[i] = C[(size_t)i * NC + C_LMAX];
    }
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    b->cheb_degree = degree;
    return TLFEA_OK;
}

namespace {
// Node planes of one instance across its long axis, from the nodes' coordinates c along it
// (nodes n0..n1-1; lo and ext the axis's minimum and extent over them, emax the longest
// element along it): the distinct coordinates when the nodes lie on few of them (a
// structured mesh), else equal-width bins of width w >= emax / 2, so that an element's
// interval of length <= 2w meets at most three consecutive bins (W <= 2).  Returns the plane
// count; plane[n - n0] receives the local plane of node n.
int instance_planes(const double* c, double lo, double ext, double emax, int n0, int n1, std::vector<int>& plane) {
    const int nn = n1 - n0;
    plane.assign(nn, 0);
    if (!(ext > 0.0)) return 1;
    std::vector<std::pair<double, int>> v(nn);
    for (int n = n0; n < n1; n++) v[n - n0] = {c[n], n - n0};
    std::sort(v.begin(), v.end());
    const double tol = 1e-9 * ext;
    int np = 0;
    for (int k = 0; k < nn; k++) {
        if (k > 0 && v[k].first - v[k - 1].first > tol) np++;
        plane[v[k].second] = np;
    }
    np++;
    if (np * 4 <= nn) return np;  // lattice planes
    const double w = std::max(0.5 * emax * (1.0 + 1e-9), ext * 1e-9);
    np = std::max(1, std::min(nn, (int)std::ceil(ext / w)));
    for (int n = n0; n < n1; n++) plane[n - n0] = std::min(np - 1, (int)((c[n] - lo) / w));
    return np;
}
}  // namespace

namespace {
// Instances whose coarse matrix is applied as a dense inverse: the sweep's chain costs about
// 1 us per block (forward and backward) whatever the instance's size, while the dense
// product streams (n_c)^2 doubles; instances are taken longest chain first while that
// lowers the modelled time of the coarse solve (the longest remaining chain against the
// dense bytes at ~4 TB/s), within 4 GiB, and only for a clear (> 20%) gain.  Each chosen
// inverse is formed column by column by the same sweep, all columns in parallel.
int coarse_dense_select(tlfea_batch* b, hipStream_t s) {
    const int NB = b->cnb, n = b->n_inst;
    const std::vector<int>& inst_cblk = b->h_inst_cblk;
    const double T_BLOCK = 1.0e-6 * std::max(1.0, NB / 12.0), BW = 4.0e12;
    constexpr double BUDGET = 4.0 * (1 << 30);
    b->h_cdense_off.assign(n, -1);
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    auto nbk = [&](int i) { return inst_cblk[i + 1] - inst_cblk[i]; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return nbk(a) > nbk(c); });
    const double sweep_only = nbk(order[0]) * T_BLOCK;
    double best = sweep_only, bytes = 0.0;
    int best_m = 0;
    for (int m = 1; m <= n; m++) {
        const double nc = (double)nbk(order[m - 1]) * NB;
        bytes += nc * nc * sizeof(double);
        if (bytes > BUDGET) break;
        const double cost = bytes / BW + (m < n ? nbk(order[m]) * T_BLOCK : 0.0);
        if (cost < best) {
            best = cost;
            best_m = m;
        }
    }
    if (best_m == 0 || best > 0.8 * sweep_only) return TLFEA_OK;
    std::vector<int> rows, rinst;
    int64_t total = 0;
    for (int j = 0; j < best_m; j++) {
        const int i = order[j];
        const int64_t nc = (int64_t)nbk(i) * NB;
        b->h_cdense_off[i] = total;
        total += nc * nc;
        for (int64_t r = 0; r < nc; r++) {
            rows.push_back((int)((int64_t)inst_cblk[i] * NB + r));
            rinst.push_back(i);
        }
    }
    int rc;
    if ((rc = dalloc(&b->d_cdense, (size_t)total)) || (rc = dalloc(&b->d_cdense_off, (size_t)n)) ||
        (rc = dalloc(&b->d_cd_row, rows.size())) || (rc = dalloc(&b->d_cd_inst, rinst.size())))
        return rc;
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_cdense_off, b->h_cdense_off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                                 s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_cd_row, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_cd_inst, rinst.data(), sizeof(int) * rinst.size(), hipMemcpyHostToDevice, s));
    // the host tables above stay alive until the copies have run
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    b->n_cd_rows = (int)rows.size();
    b->n_cd_inst = best_m;
    return TLFEA_OK;
}

// The chosen instances' A_c^-1, column by column (after the factorisation).
int coarse_dense_inverses(tlfea_batch* b, hipStream_t s) {
    for (int i = 0; i < b->n_inst; i++) {
        if (b->h_cdense_off.empty() || b->h_cdense_off[i] < 0) continue;
        if (int rc = launch_coarse_inverse(b->cnb, b->h_inst_cblk[i], b->h_inst_cblk[i + 1], b->d_cdiag, b->d_clow,
                                           b->d_cg, b->d_cdense + b->h_cdense_off[i], s))
            return rc;
    }
    return TLFEA_OK;
}
}  // namespace

namespace {
// The geometric part of the coarse space, built once per batch (the rest positions and
// fixed DOFs do not change after creation): node planes, rigid-mode offsets, plane tables,
// the block layout and the choice of dense inverses, with the device arrays they need.
int coarse_geometry(tlfea_batch* b, hipStream_t s) {
    const int N = b->n_nodes, NE = b->n_elem, nn = b->nn, n_inst = b->n_inst;
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    // device scratch: the CG workspace (free between solves) for the per-node arrays, and a
    // few per-instance arrays freed on return
    double* d_cax = b->d_r;
    unsigned char* d_fixed = reinterpret_cast<unsigned char*>(b->d_p);
    int* d_ninst = reinterpret_cast<int*>(b->d_q);
    int* d_local = reinterpret_cast<int*>(b->d_z);
    struct Scratch {
        double* bbox = nullptr;
        unsigned long long* emax = nullptr;
        int *axis = nullptr, *span = nullptr, *noff = nullptr;
        double *cen = nullptr, *size = nullptr;
        ~Scratch() {
            for (void* q : {(void*)bbox, (void*)emax, (void*)axis, (void*)span, (void*)noff, (void*)cen, (void*)size})
                if (q) (void)hipFree(q);
        }
    } t;
    int rc;
    if ((rc = dalloc(&t.bbox, 6 * (size_t)n_inst)) || (rc = dalloc(&t.emax, n_inst)) ||
        (rc = dalloc(&t.axis, n_inst)) || (rc = dalloc(&t.span, n_inst)) || (rc = dalloc(&t.noff, n_inst + 1)))
        return rc;
    // per-instance bounding boxes -> long axis, then the nodes' coordinate along it, their
    // fixed flags and the longest element along it
    TLFEA_HIP_TRY(hipMemcpyAsync(t.noff, b->h_node_off.data(), sizeof(int) * (n_inst + 1), hipMemcpyHostToDevice, s));
    k_elem_inst<<<n_inst, 256, 0, s>>>(n_inst, t.noff, d_ninst);
    k_geom_bbox<<<n_inst, 256, 0, s>>>(t.noff, b->d_X, b->d_Y, b->d_Z, t.bbox);
    std::vector<double> bbox(6 * (size_t)n_inst);
    TLFEA_HIP_TRY(hipMemcpyAsync(bbox.data(), t.bbox, sizeof(double) * bbox.size(), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    std::vector<int> inst_ax(n_inst, 0);
    for (int i = 0; i < n_inst; i++)
        for (int d = 1; d < 3; d++)
            if (bbox[6 * i + 3 + d] - bbox[6 * i + d] > bbox[6 * i + 3 + inst_ax[i]] - bbox[6 * i + inst_ax[i]])
                inst_ax[i] = d;
    TLFEA_HIP_TRY(hipMemcpyAsync(t.axis, inst_ax.data(), sizeof(int) * n_inst, hipMemcpyHostToDevice, s));
    k_geom_axis<<<grid_for(N, 256), 256, 0, s>>>(N, d_ninst, t.axis, b->d_X, b->d_Y, b->d_Z, b->d_mask, d_cax,
                                                 d_fixed);
    TLFEA_HIP_TRY(hipMemsetAsync(t.emax, 0, sizeof(unsigned long long) * n_inst, s));
    k_geom_elem_extent<<<grid_for(NE, 256), 256, 0, s>>>(NE, nn, b->d_conn, b->d_elem_inst, d_cax, t.emax);
    TLFEA_LAUNCH_CHECK();
    std::unique_ptr<double[]> cax(new double[N]);
    std::unique_ptr<unsigned char[]> fixed(new unsigned char[N]);
    std::vector<unsigned long long> emax_bits(n_inst);
    TLFEA_HIP_TRY(hipMemcpyAsync(cax.get(), d_cax, sizeof(double) * N, hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(fixed.get(), d_fixed, N, hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(emax_bits.data(), t.emax, sizeof(unsigned long long) * n_inst, hipMemcpyDeviceToHost,
                                 s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    // instances on host threads: each owns its node, element and plane ranges
    const int nthreads = std::max(1, std::min<int>(n_inst, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    auto parallel_instances = [&](auto&& fn) {
        std::vector<std::thread> pool;
        for (int th = 0; th < nthreads; th++)
            pool.emplace_back([&, th] {
                for (int i = (int)((int64_t)n_inst * th / nthreads); i < (int)((int64_t)n_inst * (th + 1) / nthreads); i++)
                    fn(i);
            });
        for (auto& th : pool) th.join();
    };
    // planes per instance (host: one sort of the axis coordinate), then the widest element
    // span W over them (device)
    std::unique_ptr<int[]> local(new int[N]);
    std::vector<int> inst_np(n_inst), inst_w(n_inst, 1);
    parallel_instances([&](int i) {
        const int n0 = b->h_node_off[i], n1 = b->h_node_off[i + 1], ax = inst_ax[i];
        double emax;
        std::memcpy(&emax, &emax_bits[i], sizeof(double));
        std::vector<int> pl;
        inst_np[i] = instance_planes(cax.get(), bbox[6 * i + ax], bbox[6 * i + 3 + ax] - bbox[6 * i + ax], emax, n0, n1,
                                     pl);
        for (int n = n0; n < n1; n++) local[n] = pl[n - n0];
    });
    TLFEA_HIP_TRY(hipMemcpyAsync(d_local, local.get(), sizeof(int) * N, hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemsetAsync(t.span, 0, sizeof(int) * n_inst, s));
    k_geom_elem_span<<<grid_for(NE, 256), 256, 0, s>>>(NE, nn, b->d_conn, b->d_elem_inst, d_local, t.span);
    TLFEA_LAUNCH_CHECK();
    std::vector<int> span(n_inst);
    TLFEA_HIP_TRY(hipMemcpyAsync(span.data(), t.span, sizeof(int) * n_inst, hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < n_inst; i++) inst_w[i] = std::max(1, span[i]);
    const int W = *std::max_element(inst_w.begin(), inst_w.end());
    if (W > 3) {
        set_error("coarse correction: elements span %d node planes (at most 3 supported)", W);
        return TLFEA_ECAPACITY;
    }
    // the in-plane modes where the blocks stay within CNB_MAX
    const int cm = (b->cm_req == CM_MAX && CM_MAX * W <= CNB_MAX) ? CM_MAX : 6;
    const int NB = cm * W;
    // global planes, padded per instance to a multiple of W
    std::vector<int> inst_poff(n_inst + 1, 0), inst_pn(n_inst);
    for (int i = 0; i < n_inst; i++) {
        inst_pn[i] = (inst_np[i] + W - 1) / W * W;
        inst_poff[i + 1] = inst_poff[i] + inst_pn[i];
    }
    const int P = inst_poff[n_inst];
    std::vector<int> plane_inst(P), plane_local(P), plane_axis(P), inst_cblk(n_inst + 1), cnt(P + 1, 0);
    inst_cblk[n_inst] = P / W;
    // the free nodes of each plane (CSR)
    std::unique_ptr<int[]> node_plane(new int[N]);
    parallel_instances([&](int i) {
        inst_cblk[i] = inst_poff[i] / W;
        for (int q = 0; q < inst_pn[i]; q++) {
            plane_inst[inst_poff[i] + q] = i;
            plane_local[inst_poff[i] + q] = q;
            plane_axis[inst_poff[i] + q] = inst_ax[i];
        }
        for (int n = b->h_node_off[i]; n < b->h_node_off[i + 1]; n++) {
            node_plane[n] = fixed[n] ? -1 : inst_poff[i] + local[n];
            if (node_plane[n] >= 0) cnt[node_plane[n] + 1]++;
        }
    });
    for (int g = 0; g < P; g++) cnt[g + 1] += cnt[g];
    std::vector<int> plane_nodes(std::max(cnt[P], 1));
    parallel_instances([&](int i) {
        std::vector<int> cur(cnt.begin() + inst_poff[i], cnt.begin() + inst_poff[i + 1]);
        for (int n = b->h_node_off[i]; n < b->h_node_off[i + 1]; n++)
            if (node_plane[n] >= 0) plane_nodes[cur[node_plane[n] - inst_poff[i]]++] = n;
    });
    // device tables
    for (void* q : {(void*)b->d_node_plane, (void*)b->d_node_r, (void*)b->d_plane_off, (void*)b->d_plane_nodes,
                    (void*)b->d_plane_inst, (void*)b->d_plane_local, (void*)b->d_inst_cblk, (void*)b->d_cdiag,
                    (void*)b->d_clow, (void*)b->d_cg, (void*)b->d_crhs, (void*)b->d_cx, (void*)b->d_cdense_off,
                    (void*)b->d_cdense, (void*)b->d_cd_row, (void*)b->d_cd_inst, (void*)b->d_cinst_pn,
                    (void*)b->d_plane_axis})
        if (q) (void)hipFree(q);
    b->d_plane_axis = nullptr;
    b->d_cinst_pn = nullptr;
    b->d_cdense_off = nullptr;
    b->d_cdense = nullptr;
    b->d_cd_row = b->d_cd_inst = nullptr;
    b->n_cd_rows = b->n_cd_inst = 0;
    b->d_node_plane = b->d_plane_off = b->d_plane_nodes = b->d_plane_inst = b->d_plane_local = b->d_inst_cblk =
        nullptr;
    b->d_node_r = b->d_cdiag = b->d_clow = b->d_cg = b->d_crhs = b->d_cx = nullptr;
    const size_t nblk = (size_t)P / W, bsz = (size_t)NB * NB;
    int* d_inst_pn = nullptr;
    if ((rc = dalloc(&b->d_node_plane, N)) || (rc = dalloc(&b->d_node_r, 3 * (size_t)N)) ||
        (rc = dalloc(&b->d_plane_off, P + 1)) || (rc = dalloc(&b->d_plane_nodes, plane_nodes.size())) ||
        (rc = dalloc(&b->d_plane_inst, P)) || (rc = dalloc(&b->d_plane_local, P)) ||
        (rc = dalloc(&b->d_inst_cblk, n_inst + 1)) || (rc = dalloc(&b->d_cdiag, nblk * bsz)) ||
        (rc = dalloc(&b->d_clow, nblk * bsz)) || (rc = dalloc(&b->d_cg, nblk * bsz)) ||
        (rc = dalloc(&b->d_crhs, cm * (size_t)P)) || (rc = dalloc(&b->d_cx, cm * (size_t)P)) ||
        (rc = dalloc(&b->d_plane_axis, P)) || (rc = dalloc(&d_inst_pn, n_inst)) || (rc = dalloc(&t.cen, 3 * (size_t)P)) ||
        (rc = dalloc(&t.size, P)))
        return rc;
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_node_plane, node_plane.get(), sizeof(int) * N, hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_plane_off, cnt.data(), sizeof(int) * (P + 1), hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_plane_nodes, plane_nodes.data(), sizeof(int) * plane_nodes.size(),
                                 hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_plane_inst, plane_inst.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_plane_local, plane_local.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_plane_axis, plane_axis.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(
        hipMemcpyAsync(b->d_inst_cblk, inst_cblk.data(), sizeof(int) * (n_inst + 1), hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(d_inst_pn, inst_pn.data(), sizeof(int) * n_inst, hipMemcpyHostToDevice, s));
    // plane centroids and sizes, then every node's scaled offset r = (X - c) / size
    k_geom_plane_frame<<<P, 256, 0, s>>>(b->d_plane_off, b->d_plane_nodes, b->d_X, b->d_Y, b->d_Z, t.cen, t.size);
    k_geom_node_r<<<grid_for(N, 256), 256, 0, s>>>(N, b->d_node_plane, b->d_X, b->d_Y, b->d_Z, t.cen, t.size,
                                                   b->d_node_r);
    TLFEA_LAUNCH_CHECK();
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    b->cw = W;
    b->cm = cm;
    b->cnb = NB;
    b->n_planes = P;
    b->n_cblocks = (int)nblk;
    b->d_cinst_pn = d_inst_pn;
    b->h_inst_cblk = inst_cblk;
    b->restrict_wide = P > 0 && cnt[P] >= 512 * (int64_t)P;
    if ((rc = coarse_dense_select(b, s))) return rc;
    b->coarse_geom = true;
    return TLFEA_OK;
}

// The operator-dependent part, every call: A_c = P^T A P by probing, its block
// factorisation and the dense inverses.
int coarse_numeric(tlfea_batch* b, hipStream_t s) {
    const int N = b->n_nodes, W = b->cw, NB = b->cnb, P = b->n_planes;
    const size_t bsz = (size_t)NB * NB, nblk = (size_t)b->n_cblocks;
    int rc = TLFEA_OK;
    TLFEA_HIP_TRY(hipMemsetAsync(b->d_cdiag, 0, sizeof(double) * nblk * bsz, s));
    TLFEA_HIP_TRY(hipMemsetAsync(b->d_clow, 0, sizeof(double) * nblk * bsz, s));
    TLFEA_HIP_TRY(hipMemsetAsync(b->d_cx, 0, sizeof(double) * b->cm * (size_t)P, s));
    // P^T A P (A = K + c M when a mass term is set) by probing: planes coloured by local
    // index mod 2W+1 are further apart than any element reaches, so each coloured rigid
    // mode's image restricted to a plane belongs to the one coloured plane within W of it
    const int period = 2 * W + 1;
    for (int color = 0; color < period && rc == TLFEA_OK; color++)
        for (int j = 0; j < b->cm && rc == TLFEA_OK; j++) {
            k_coarse_probe_vec<<<grid_for(N, 256), 256, 0, s>>>(N, b->d_node_plane, b->d_node_r, b->d_plane_local,
                                                                b->d_plane_axis, period, color, j, b->d_p);
            if ((rc = apply64(b, b->d_p, b->d_q, nullptr, s))) break;
            // four waves per plane on large planes (as the restriction)
            (b->restrict_wide ? k_coarse_probe_store<256> : k_coarse_probe_store<64>)<<<P, b->restrict_wide ? 256 : 64,
                                                                                         0, s>>>(
                W, b->cm, period, color, j, b->d_plane_off, b->d_plane_nodes, b->d_plane_local, b->d_plane_inst,
                b->d_plane_axis, b->d_cinst_pn, b->d_node_r, b->d_q, b->d_cdiag, b->d_clow);
            if (hipGetLastError() != hipSuccess) {
                set_error("coarse probe launch failed");
                rc = TLFEA_EHIP;
            }
        }
    if (rc == TLFEA_OK) {
        if (NB == 6)
            k_coarse_factor<6><<<b->n_inst, 64, 0, s>>>(b->d_inst_cblk, b->d_cdiag, b->d_clow, b->d_cg);
        else if (NB == 11)
            k_coarse_factor<11><<<b->n_inst, 64, 0, s>>>(b->d_inst_cblk, b->d_cdiag, b->d_clow, b->d_cg);
        else if (NB == 12)
            k_coarse_factor<12><<<b->n_inst, 64, 0, s>>>(b->d_inst_cblk, b->d_cdiag, b->d_clow, b->d_cg);
        else if (NB == 22)
            k_coarse_factor_lds<22><<<b->n_inst, FACTOR_NT, 0, s>>>(b->d_inst_cblk, b->d_cdiag, b->d_clow, b->d_cg);
        else
            k_coarse_factor_lds<18><<<b->n_inst, FACTOR_NT, 0, s>>>(b->d_inst_cblk, b->d_cdiag, b->d_clow, b->d_cg);
        if (hipGetLastError() != hipSuccess) {
            set_error("coarse factor launch failed");
            rc = TLFEA_EHIP;
        }
    }
    if (rc == TLFEA_OK) rc = coarse_dense_inverses(b, s);
    const hipError_t es = hipStreamSynchronize(s);
    if (rc) return rc;
    TLFEA_HIP_TRY(es);
    return TLFEA_OK;
}
}  // namespace

extern "C" int tlfea_batch_coarse_setup(tlfea_batch* b, int enable, int* n_coarse, void* stream) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    drop_graph(b);
    full_lists(b);
    b->coarse_on = false;
    if (!enable) return TLFEA_OK;
    TLFEA_REQUIRE(b->cheb_degree > 0, "the coarse correction runs inside the Chebyshev stage: set that up first");
    hipStream_t s = as_stream(stream);
    if (!b->coarse_geom) {
        TLFEA_HIP_TRY(hipStreamSynchronize(s));
        if (int rc = coarse_geometry(b, s)) return rc;
    }
    if (int rc = coarse_numeric(b, s)) return rc;
    if (n_coarse) *n_coarse = b->cm * b->n_planes;
    b->coarse_on = true;
    return TLFEA_OK;
}

extern "C" int tlfea_batch_set_coarse_modes(tlfea_batch* b, int modes) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    TLFEA_REQUIRE(modes == 6 || modes == CM_MAX, "coarse modes per plane must be 6 or %d", CM_MAX);
    drop_graph(b);
    b->cm_req = modes;
    b->coarse_on = false;
    b->coarse_geom = false;  // rebuilt (with the new block size) by the next set-up
    return TLFEA_OK;
}

extern "C" int tlfea_batch_coarse_apply(tlfea_batch* b, const double* r, double* z, void* stream) {
    TLFEA_REQUIRE(b && r && z, "null argument");
    TLFEA_REQUIRE(b->coarse_on, "the coarse correction is not set up");
    hipStream_t s = as_stream(stream);
    k_bclear_flags<<<grid_for(b->n_inst, 256), 256, 0, s>>>(b->n_inst, b->d_i);
    launch_restrict(b, b->d_plane_off, b->d_plane_nodes, b->d_plane_inst, b->d_node_r, b->d_i, r, b->d_crhs, s);
    if (int rc = coarse_solve_all(b, s)) return rc;
    const int64_t n3 = 3 * (int64_t)b->n_nodes;
    k_coarse_prolong64<<<grid_for(n3, 256), 256, 0, s>>>(n3, CoarseProlong{b->d_node_plane, b->d_node_r, b->d_cx,
                                                                            b->cm, b->d_plane_axis},
                                                         z);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// u += z, FP64 (z from the Chebyshev stage)
__global__ __launch_bounds__(BT) void k_badd(Chunks ch, const double* __restrict__ z, double* __restrict__ u) {
    const int c = chunk_id(ch);
    for (int64_t k = 3 * (int64_t)ch.beg[c] + threadIdx.x; k < 3 * (int64_t)ch.end[c]; k += BT) u[k] += z[k];
}

extern "C" int tlfea_batch_smooth(tlfea_batch* b, const double* f, double* u, int sweeps, void* stream) {
    TLFEA_REQUIRE(b && f && u, "null argument");
    TLFEA_REQUIRE(b->cheb_degree > 0, "the Chebyshev stage is not set up");
    TLFEA_REQUIRE(sweeps >= 0, "negative sweep count");
    hipStream_t s = as_stream(stream);
    full_lists(b);
    const Chunks ch = chunks(b);
    // every instance takes part: clear the done flags the Chebyshev kernels test
    k_bclear_flags<<<grid_for(b->n_inst, 256), 256, 0, s>>>(b->n_inst, b->d_i);
    for (int k = 0; k < sweeps; k++) {
        if (int rc = apply64(b, u, b->d_q, nullptr, s)) return rc;
        k_binit<<<b->n_chunks, BT, 0, s>>>(ch, f, b->d_q, b->d_mask, b->d_r, b->d_part);
        TLFEA_HIP_TRY(hipMemcpyAsync(b->d_rold, b->d_r, sizeof(double) * 3 * (size_t)b->n_nodes,
                                     hipMemcpyDeviceToDevice, s));
        if (int rc = precondition(b, s)) return rc;
        k_badd<<<b->n_chunks, BT, 0, s>>>(ch, b->d_z, u);
        TLFEA_LAUNCH_CHECK();
    }
    return TLFEA_OK;
}

extern "C" int tlfea_batch_tangent_apply(tlfea_batch* b, const double* p, double* out, void* stream) {
    TLFEA_REQUIRE(b && p && out, "null argument");
    full_lists(b);
    return apply64(b, p, out, nullptr, as_stream(stream), 1.0, false);
}

extern "C" int tlfea_batch_tangent_apply_f32(tlfea_batch* b, const float* p, float* out, void* stream) {
    TLFEA_REQUIRE(b && p && out, "null argument");
    full_lists(b);
    return apply32(b, p, out, nullptr, as_stream(stream));
}

extern "C" int tlfea_batch_set_mass(tlfea_batch* b, const double* rho_coef, int rule) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    drop_graph(b);
    TLFEA_REQUIRE(!rho_coef || rule == 0 || rule == 1, "mass rule must be 0 (quadrature) or 1 (exact)");
    // the preconditioner was set up for the previous operator
    b->cheb_degree = 0;
    b->coarse_on = false;
    if (!rho_coef) {
        b->mass_on = false;
        return TLFEA_OK;
    }
    if (!b->d_mc && (dalloc(&b->d_mc, b->n_inst) || dalloc(&b->d_mc32, b->n_inst))) return TLFEA_EHIP;
    TLFEA_HIP_TRY(hipMemcpy(b->d_mc, rho_coef, sizeof(double) * b->n_inst, hipMemcpyHostToDevice));
    std::vector<float> mc32(rho_coef, rho_coef + b->n_inst);
    TLFEA_HIP_TRY(hipMemcpy(b->d_mc32, mc32.data(), sizeof(float) * b->n_inst, hipMemcpyHostToDevice));
    b->mass_on = true;
    b->mass_exact = rule;
    return TLFEA_OK;
}

extern "C" int tlfea_batch_set_displacement(tlfea_batch* b, const double* ux, const double* uy, const double* uz,
                                             void* stream) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    drop_graph(b);
    if (!ux && !uy && !uz) {
        b->moved = false;
        return TLFEA_OK;
    }
    TLFEA_REQUIRE(ux && uy && uz, "the displacement needs all three arrays");
    if (!b->d_ux && (dalloc(&b->d_ux, b->n_nodes) || dalloc(&b->d_uy, b->n_nodes) || dalloc(&b->d_uz, b->n_nodes)))
        return TLFEA_EHIP;
    hipStream_t s = as_stream(stream);
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_ux, ux, sizeof(double) * b->n_nodes, hipMemcpyDeviceToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_uy, uy, sizeof(double) * b->n_nodes, hipMemcpyDeviceToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_uz, uz, sizeof(double) * b->n_nodes, hipMemcpyDeviceToDevice, s));
    b->moved = true;
    return TLFEA_OK;
}

extern "C" int tlfea_batch_internal_force(tlfea_batch* b, double* f, void* stream) {
    TLFEA_REQUIRE(b && f, "null argument");
    if ((size_t)b->plan->max_u * 9 * sizeof(double) > 160 * 1024) {
        set_error("a block touches %d nodes: beyond the LDS budget of the rest-coordinate force", b->plan->max_u);
        return TLFEA_ECAPACITY;
    }
    const tlfea_material none{0.0, 0.0, 0.0, 0.0};
    return launch_force_rest(b->plan, pos_x(b), pos_y(b), pos_z(b), b->d_X, b->d_Y, b->d_Z, b->quad, none, nullptr, f,
                             0, nullptr, as_stream(stream), b->d_elem_inst, b->d_lam, b->d_mu, b->moved);
}

extern "C" int tlfea_batch_dot(tlfea_batch* b, const double* a, const double* c, double* out, void* stream) {
    TLFEA_REQUIRE(b && a && c && out, "null argument");
    hipStream_t s = as_stream(stream);
    const Chunks all{b->d_chunk_inst, b->d_chunk_beg, b->d_chunk_end, nullptr};
    k_bdot_ac<<<b->n_chunks, BT, 0, s>>>(all, a, c, b->d_part);
    k_binst_sum<<<b->n_inst, 64, 0, s>>>(b->n_inst, b->d_inst_chunk_off, b->d_part, out);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_batch_apply(tlfea_batch* b, double k_scale, int with_mass, int masked, const double* p,
                                 double* out, void* stream) {
    TLFEA_REQUIRE(b && p && out, "null argument");
    TLFEA_REQUIRE(!with_mass || b->mass_on, "no mass coefficients set");
    full_lists(b);
    return apply64(b, p, out, nullptr, as_stream(stream), k_scale, with_mass != 0, masked != 0);
}

extern "C" int tlfea_batch_set_compaction(tlfea_batch* b, double ratio) {
    TLFEA_REQUIRE(b != nullptr, "null batch");
    TLFEA_REQUIRE(ratio >= 0.0 && ratio <= 1.0, "compaction ratio must be in [0, 1]");
    b->compact_ratio = ratio;
    return TLFEA_OK;
}

extern "C" int tlfea_batch_cg_solve(tlfea_batch* b, const double* f, double* u_out, int use_guess, double tol,
                                    int max_iter, int poll, int use_graph, int precond, int* iters, double* rel_res,
                                    int* status, int* total_iters, void* stream) {
    TLFEA_REQUIRE(b && f && u_out, "null argument");
    TLFEA_REQUIRE(precond == 0 || b->cheb_degree > 0, "preconditioner requested but not set up");
    TLFEA_REQUIRE(max_iter >= 0 && tol >= 0.0, "bad tolerance or iteration limit");
    if (poll <= 0) poll = b->poll_default;
    // the single-reduction iteration keys its parity slots to the iteration index, which
    // a replayed graph of an even number of iterations keeps aligned
    if (!precond && (poll & 1)) poll += 1;
    hipStream_t caller = as_stream(stream);
    hipStream_t s = b->s_own;
    const double compact_ratio = b->compact_ratio;
    // order the solve after the caller's pending work
    TLFEA_HIP_TRY(hipStreamSynchronize(caller));
    full_lists(b);
    const Chunks ch = chunks(b);
    const size_t nd = 3 * (size_t)b->n_nodes;
    double* u = b->d_usol;
    if (use_guess)
        TLFEA_HIP_TRY(hipMemcpyAsync(u, u_out, sizeof(double) * nd, hipMemcpyDeviceToDevice, s));
    else
        TLFEA_HIP_TRY(hipMemsetAsync(u, 0, sizeof(double) * nd, s));
    // r = f - K u0
    if (int rc = apply64(b, u, b->d_q, nullptr, s)) return rc;
    k_binit<<<b->n_chunks, BT, 0, s>>>(ch, f, b->d_q, b->d_mask, b->d_r, b->d_part, b->d_all);
    k_breduce_init<<<b->n_inst, 64, 0, s>>>(b->n_inst, b->d_inst_chunk_off, b->d_part, b->d_s, b->d_i, b->d_all, tol);
    k_bzero_u<<<b->n_chunks, BT, 0, s>>>(ch, b->d_i, u);
    if (precond) {
        // z0 = M^-1 r0, rz = r0.z0, p = z0 (r_old = r0 makes the z.r_old partial harmless)
        TLFEA_HIP_TRY(hipMemcpyAsync(b->d_rold, b->d_r, sizeof(double) * nd, hipMemcpyDeviceToDevice, s));
        if (int rc = precondition(b, s)) return rc;
        k_breduce_rz0<<<insts(b).n, 64, 0, s>>>(insts(b), b->d_inst_chunk_off, b->d_part, b->d_s, b->d_i);
        k_bdir<<<n_chunk_launch(b), BT, 0, s>>>(ch, b->d_p, b->d_z, b->d_s, b->d_i);
    } else {
        // p = s = 0 and fresh parity slots: the first iteration then takes p = r
        TLFEA_HIP_TRY(hipMemsetAsync(b->d_p, 0, sizeof(double) * nd, s));
        TLFEA_HIP_TRY(hipMemsetAsync(b->d_z, 0, sizeof(double) * nd, s));
        k_bcg_init_slots<<<grid_for(b->n_inst, 256), 256, 0, s>>>(b->n_inst, b->d_s);
    }
    TLFEA_LAUNCH_CHECK();
    int active = 0;
    TLFEA_HIP_TRY(hipMemcpyAsync(&active, b->d_all, sizeof(int), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    int it = 0;
    std::vector<int> hflags((size_t)b->n_inst * NI), done(b->n_inst, 0);
    auto iteration = [&](hipStream_t st, int k) {
        return precond ? pcg_iteration(b, u, tol, st) : cg_iteration(b, u, tol, st, k & 1);
    };
    if (use_graph && (b->graph == nullptr || b->graph_iters != poll || b->graph_precond != precond ||
                      b->graph_tol != tol))
        drop_graph(b);
    while (active > 0 && it < max_iter) {
        const int n = std::min(poll, max_iter - it);
        if (use_graph && n == poll) {
            if (!b->graph) {
                hipGraph_t g;
                TLFEA_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int rc = TLFEA_OK;
                for (int k = 0; k < poll && rc == TLFEA_OK; k++) rc = iteration(s, k);  // it % poll == 0
                const hipError_t ec = hipStreamEndCapture(s, &g);
                if (rc) return rc;
                if (ec != hipSuccess) {
                    set_error("graph capture failed: %s", hipGetErrorString(ec));
                    return TLFEA_EHIP;
                }
                const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    set_error("graph instantiation failed: %s", hipGetErrorString(ei));
                    b->graph = nullptr;
                    return TLFEA_EHIP;
                }
                b->graph_iters = poll;
                b->graph_precond = precond;
                b->graph_tol = tol;
            }
            TLFEA_HIP_TRY(hipGraphLaunch(b->graph, s));
        } else {
            for (int k = 0; k < n; k++)
                if (int rc = iteration(s, it + k)) return rc;
        }
        it += n;
        TLFEA_HIP_TRY(hipMemcpyAsync(&active, b->d_all, sizeof(int), hipMemcpyDeviceToHost, s));
        TLFEA_HIP_TRY(hipStreamSynchronize(s));
        if (active > 0 && compact_ratio > 0.0) {
            // drop finished instances from the launches once they are a large enough share
            TLFEA_HIP_TRY(hipMemcpy(hflags.data(), b->d_i, sizeof(int) * hflags.size(), hipMemcpyDeviceToHost));
            int64_t act_elems = 0;
            for (int i = 0; i < b->n_inst; i++) {
                done[i] = hflags[(size_t)i * NI + I_DONE];
                if (!done[i]) act_elems += b->h_elem_off[i + 1] - b->h_elem_off[i];
            }
            if ((double)act_elems < compact_ratio * (double)b->launched_elems)
                if (int rc = compact_lists(b, done)) return rc;
        }
    }
    full_lists(b);
    if (!precond && active > 0) {
        // out of iterations: the residual of the last update has not been tested yet
        k_bdot_rr<<<b->n_chunks, BT, 0, s>>>(chunks(b), b->d_r, b->d_part);
        k_breduce_tail<<<b->n_inst, 64, 0, s>>>(b->n_inst, b->d_inst_chunk_off, b->d_part, b->d_s, b->d_i, tol);
        TLFEA_LAUNCH_CHECK();
    }
    TLFEA_HIP_TRY(hipMemcpyAsync(u_out, u, sizeof(double) * nd, hipMemcpyDeviceToDevice, s));
    // s is a non-blocking stream, unordered with the null stream: the copies of the
    // per-instance results go on s, after the tail kernels that write them
    std::vector<double> S((size_t)b->n_inst * NS);
    std::vector<int> I((size_t)b->n_inst * NI);
    TLFEA_HIP_TRY(hipMemcpyAsync(S.data(), b->d_s, sizeof(double) * S.size(), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(I.data(), b->d_i, sizeof(int) * I.size(), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    int bad = 0;
    for (int i = 0; i < b->n_inst; i++) {
        if (iters) iters[i] = I[(size_t)i * NI + I_ITERS];
        if (rel_res) rel_res[i] = S[(size_t)i * NS + S_REL];
        int st = TLFEA_OK;
        if (I[(size_t)i * NI + I_NOTPD])
            st = TLFEA_ENOTPD;
        else if (!I[(size_t)i * NI + I_DONE])
            st = TLFEA_ENOCONV;
        if (status) status[i] = st;
        bad += st != TLFEA_OK;
    }
    if (total_iters) *total_iters = it;
    if (bad) {
        set_error("%d of %d instances did not converge", bad, b->n_inst);
        return TLFEA_ENOCONV;
    }
    return TLFEA_OK;
}

// ---- batched SyncedNesterov ---------------------------------------------------------------
// One implicit step per instance, g(v) = M (v - v_n) / dt + f_int(X + u_n + dt v) - f_ext = 0
// (SyncedNesterov.cu:32-95), minimised by the reference's inner Nesterov loop
// (SyncedNesterov.cu:97-360) for every instance at once, each with its own step size,
// momentum sequence and stopping test.  Fixed DOFs are eliminated (v = 0 there) instead of
// carried by the augmented-Lagrangian outer loop: with no constraint left that loop ends
// after its first pass.
namespace tlfea_hip {

// per-instance Nesterov scalars: t_k and t_(k+1) of the momentum sequence, |g| of the
// previous iteration, |g| at the first iteration, |g| now, |g| / |g_0|, the step size
enum { N_T = 0, N_T1, N_GPREV, N_G0, N_G, N_REL, N_ALPHA, NSN };

__global__ void k_bnes_init(int n_inst, const double* __restrict__ alpha, double* __restrict__ ns,
                            int* __restrict__ I, int* __restrict__ rs, int* __restrict__ all) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) all[0] = n_inst;
    if (i >= n_inst) return;
    double* s = ns + (int64_t)i * NSN;
    s[N_T] = 1.0;
    s[N_T1] = 0.5 * (1.0 + sqrt(5.0));
    s[N_GPREV] = s[N_G0] = s[N_G] = s[N_REL] = 0.0;
    s[N_ALPHA] = alpha[i];
    int* fl = I + (int64_t)i * NI;
    fl[I_DONE] = fl[I_ITERS] = fl[I_NOTPD] = fl[I_ZERO_F] = fl[I_RESET] = fl[I_HALVE] = 0;
    rs[i] = 0;
}


// The predict step of SyncedNesterov.cu:152-175 is formed where it is used: the force
// kernel's staging builds y, dv = y - v_n and u_n + dt y per node from (v_k, v_(k-1), v_n,
// u_n) and the instance's beta = (t_k - 1) / t_(k+1) (ForceMass), and this kernel forms
// the same y per DOF: g = (f_int + M dv / dt) - f_ext from the block partials (0 on fixed
// DOFs), v_next = y - alpha g into v, v_(k-1) <- v_k; chunk partials |g|^2, |v_next|^2,
// |v_k|^2, g.(v_next - v_k) (SyncedNesterov.cu:32-95, 177-224).  An instance whose step
// size was just halved starts again from v_k = v_(k-1) = v_n.  all[0] is zeroed for
// k_bnes_finalize's count.
__global__ __launch_bounds__(BT) void k_bnes_grad(Chunks ch, const int* __restrict__ g_off,
                                                  const int* __restrict__ g_idx, const double* __restrict__ part,
                                                  const unsigned char* __restrict__ mask,
                                                  const double* __restrict__ fext, const double* __restrict__ vn,
                                                  double* __restrict__ vkm1, double* __restrict__ v,
                                                  const double* __restrict__ ns, const int* __restrict__ I,
                                                  double* __restrict__ cpart, int* __restrict__ all) {
    if (blockIdx.x == 0 && threadIdx.x == 0) all[0] = 0;
    const int c = chunk_id(ch);
    const int inst = ch.inst[c];
    if (I[inst * NI + I_DONE]) return;
    const double alpha = ns[inst * NSN + N_ALPHA];
    const double beta = (ns[inst * NSN + N_T] - 1.0) / ns[inst * NSN + N_T1];
    const bool reset = I[inst * NI + I_RESET] != 0;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int n = ch.beg[c] + threadIdx.x; n < ch.end[c]; n += BT) {
        double s3[3] = {0.0, 0.0, 0.0};
        gather_slots3(g_idx, part, g_off[n], g_off[n + 1], s3[0], s3[1], s3[2]);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int64_t k = 3 * (int64_t)n + d;
            const bool fx = mask[k] != 0;
            const double vo = reset ? vn[k] : v[k];
            const double yv = fx ? 0.0 : vo + beta * (vo - (reset ? vo : vkm1[k]));
            const double g = fx ? 0.0 : s3[d] - fext[k];
            const double vnext = fx ? 0.0 : yv - alpha * g;
            v[k] = vnext;
            vkm1[k] = vo;
            acc[0] += g * g;
            acc[1] += vnext * vnext;
            acc[2] += vo * vo;
            acc[3] += g * (vnext - vo);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const double t = chunk_sum(acc[j]);
        if (threadIdx.x == 0) cpart[4 * (int64_t)c + j] = t;
    }
}

// Per instance: the stopping tests of SyncedNesterov.cu:226-271 (inner_tol > 0: | |g| -
// |g_prev| | or | |v_next| - |v_k| | below it after the first iteration), the relative test
// |g| <= rel_tol |g_0| (rel_tol > 0) and |g| = 0; then the momentum sequence advances, or,
// with restart, starts again (t_(k+1) = 1: no momentum in the next step) when
// g.(v_next - v_k) > 0 (O'Donoghue & Candes, Found. Comput. Math. 15 (2015) 715-732).
// Safeguard (not in the reference, whose step size is fixed): a gradient that is not
// finite or has grown past 1e4 |g_0| means the step size exceeds 2 / L along the path (the
// tangent stiffens under large deflection); the instance then halves its step size and
// starts again from v_n, at most 40 times before it is given up as not converged.
// One wave per listed instance; all[0] (zeroed by k_bnes_grad) counts the instances still
// running.
__global__ __launch_bounds__(64) void k_bnes_finalize(Insts in, const int* __restrict__ ico,
                                                      const double* __restrict__ cpart, double* __restrict__ ns,
                                                      int* __restrict__ I, int* __restrict__ rs, int* __restrict__ all,
                                                      double inner_tol, double rel_tol, int restart) {
    const int i = inst_at(in, blockIdx.x);
    int* fl = I + (int64_t)i * NI;
    if (fl[I_DONE]) return;
    // the chunk partials in a fixed order: lane l sums chunks l, l + 64, ..., then a
    // butterfly over the lanes
    double sm[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = ico[i] + (int)threadIdx.x; c < ico[i + 1]; c += 64)
#pragma unroll
        for (int k = 0; k < 4; k++) sm[k] += cpart[4 * (int64_t)c + k];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sm[k] += __shfl_xor(sm[k], o, 64);
    if (threadIdx.x == 0) {
        double* s = ns + (int64_t)i * NSN;
        const double ng = sqrt(sm[0]), nvn = sqrt(sm[1]), nvk = sqrt(sm[2]);
        const int it = fl[I_ITERS];
        fl[I_RESET] = 0;
        if (it == 0) s[N_G0] = ng;
        fl[I_ITERS] = it + 1;
        if (!(ng <= 1e4 * s[N_G0])) {  // also catches a NaN or an infinite |g|
            if (it == 0 || fl[I_HALVE] >= 40) {
                fl[I_NOTPD] = 1;
                fl[I_DONE] = 1;
                s[N_REL] = ng / s[N_G0];
                return;
            }
            s[N_ALPHA] *= 0.5;
            fl[I_HALVE] += 1;
            fl[I_RESET] = 1;
            s[N_T] = 1.0;
            s[N_T1] = 0.5 * (1.0 + sqrt(5.0));
            s[N_GPREV] = s[N_G0];
            atomicAdd(all, 1);
            return;
        }
        bool stop = ng == 0.0;
        if (it > 0 && inner_tol > 0.0 && (fabs(ng - s[N_GPREV]) < inner_tol || fabs(nvn - nvk) < inner_tol))
            stop = true;
        s[N_REL] = s[N_G0] > 0.0 ? ng / s[N_G0] : 0.0;
        if (rel_tol > 0.0 && s[N_REL] <= rel_tol) stop = true;
        s[N_GPREV] = ng;
        s[N_G] = ng;
        if (restart && sm[3] > 0.0) {
            s[N_T] = 1.0;
            rs[i] += 1;
        } else {
            s[N_T] = s[N_T1];
        }
        s[N_T1] = 0.5 * (1.0 + sqrt(1.0 + 4.0 * s[N_T] * s[N_T]));
        if (stop)
            fl[I_DONE] = 1;
        else
            atomicAdd(all, 1);
    }
}

}  // namespace tlfea_hip

namespace {
int nesterov_iteration(tlfea_batch* b, double dt, int mass_exact, const double* u_n, const double* v_n,
                       const double* f_ext, double* v, double inner_tol, double rel_tol, int restart, hipStream_t s) {
    const Chunks ch = chunks(b);
    const int nc = n_chunk_launch(b);
    ForceMass fm;
    fm.nes_v = v;
    fm.nes_vkm1 = b->d_nvk;
    fm.nes_vn = v_n;
    fm.nes_un = u_n;
    fm.nes_mask = b->d_mask;
    fm.node_inst = b->d_node_inst;
    fm.nes_s = b->d_ns + N_T;  // t_k, then t_(k+1)
    fm.nes_sstride = NSN;
    fm.nes_reset = b->d_i + I_RESET;
    fm.dt = dt;
    fm.inst_mc = b->d_nmc;
    fm.mass_exact = mass_exact;
    fm.inst_done = b->d_i + I_DONE;
    fm.done_stride = NI;
    fm.blk_list = b->compact ? b->d_blk_list : nullptr;
    fm.n_list = b->n_blk_list;
    const tlfea_material none{0.0, 0.0, 0.0, 0.0};
    if (int rc = launch_force_rest_mass_partials(b->plan, nullptr, nullptr, nullptr, b->d_X, b->d_Y, b->d_Z, b->quad,
                                                 none, b->d_elem_inst, b->d_lam, b->d_mu, fm, s))
        return rc;
    k_bnes_grad<<<nc, BT, 0, s>>>(ch, b->plan->d_g_off, b->plan->d_g_idx, b->plan->d_part, b->d_mask, f_ext, v_n,
                                  b->d_nvk, v, b->d_ns, b->d_i, b->d_npart, b->d_all);
    const Insts in = insts(b);
    k_bnes_finalize<<<in.n, 64, 0, s>>>(in, b->d_inst_chunk_off, b->d_npart, b->d_ns, b->d_i, b->d_nrs, b->d_all,
                                        inner_tol, rel_tol, restart);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}
}  // namespace

extern "C" int tlfea_batch_nesterov_solve(tlfea_batch* b, double dt, const double* rho, int mass_rule,
                                          const double* alpha, const double* u_n, const double* v_n,
                                          const double* f_ext, double* v, int use_guess, double inner_tol,
                                          double rel_tol, int max_inner, int restart, int poll, int use_graph,
                                          int* iters, double* rel, int* restarts, int* halvings, int* status,
                                          int* total_iters, void* stream) {
    TLFEA_REQUIRE(b && rho && alpha && u_n && v_n && f_ext && v, "null argument");
    TLFEA_REQUIRE(dt > 0.0, "time step must be positive");
    TLFEA_REQUIRE(mass_rule == 0 || mass_rule == 1, "mass rule must be 0 (quadrature) or 1 (exact)");
    TLFEA_REQUIRE(max_inner >= 0 && inner_tol >= 0.0 && rel_tol >= 0.0, "bad tolerance or iteration limit");
    TLFEA_REQUIRE(inner_tol > 0.0 || rel_tol > 0.0, "no stopping test: set inner_tol or rel_tol");
    TLFEA_REQUIRE(is_standard_rule(b->nn, b->quad), "the batched Nesterov force needs the standard quadrature rule");
    if ((size_t)b->plan->max_u * 12 * sizeof(double) > 160 * 1024) {
        set_error("a block touches %d nodes: beyond the LDS budget of the force with a mass term", b->plan->max_u);
        return TLFEA_ECAPACITY;
    }
    for (int i = 0; i < b->n_inst; i++)
        TLFEA_REQUIRE(alpha[i] > 0.0 && rho[i] > 0.0, "instance %d: step size and density must be positive", i);
    if (poll <= 0) poll = b->poll_default;
    hipStream_t caller = as_stream(stream);
    hipStream_t s = b->s_own;
    TLFEA_HIP_TRY(hipStreamSynchronize(caller));
    const int N = b->n_nodes;
    const size_t nd = 3 * (size_t)N;
    int rc;
    if (!b->d_nvk) {
        if ((rc = dalloc(&b->d_nvk, nd)) || (rc = dalloc(&b->d_node_inst, N)) ||
            (rc = dalloc(&b->d_ns, (size_t)b->n_inst * NSN)) || (rc = dalloc(&b->d_nmc, b->n_inst)) ||
            (rc = dalloc(&b->d_nrs, b->n_inst)))
            return rc;
        std::vector<int> ni(N);
        for (int i = 0; i < b->n_inst; i++)
            for (int n = b->h_node_off[i]; n < b->h_node_off[i + 1]; n++) ni[n] = i;
        TLFEA_HIP_TRY(hipMemcpy(b->d_node_inst, ni.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    }
    if (b->n_npart < b->n_chunks) {
        if (b->d_npart) (void)hipFree(b->d_npart);
        b->d_npart = nullptr;
        b->n_npart = 0;
        if ((rc = dalloc(&b->d_npart, 4 * (size_t)b->n_chunks))) return rc;
        b->n_npart = b->n_chunks;
    }
    std::vector<double> mc(b->n_inst);
    for (int i = 0; i < b->n_inst; i++) mc[i] = rho[i] / dt;
    std::vector<double> al(alpha, alpha + b->n_inst);
    double* d_al = nullptr;
    if ((rc = dalloc(&d_al, b->n_inst))) return rc;
    std::unique_ptr<double, void (*)(double*)> al_guard(d_al, [](double* p) { (void)hipFree(p); });
    TLFEA_HIP_TRY(hipMemcpyAsync(d_al, al.data(), sizeof(double) * b->n_inst, hipMemcpyHostToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_nmc, mc.data(), sizeof(double) * b->n_inst, hipMemcpyHostToDevice, s));
    drop_graph(b);
    full_lists(b);
    // v_k = v_(k-1) = v_guess (SyncedNesterov.cu:129-134): the caller's v, or v_n
    if (!use_guess) TLFEA_HIP_TRY(hipMemcpyAsync(v, v_n, sizeof(double) * nd, hipMemcpyDeviceToDevice, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(b->d_nvk, v, sizeof(double) * nd, hipMemcpyDeviceToDevice, s));
    k_bnes_init<<<grid_for(b->n_inst, 256), 256, 0, s>>>(b->n_inst, d_al, b->d_ns, b->d_i, b->d_nrs, b->d_all);
    TLFEA_LAUNCH_CHECK();
    int active = b->n_inst, it = 0;
    std::vector<int> hflags((size_t)b->n_inst * NI), done(b->n_inst, 0);
    rc = TLFEA_OK;
    while (rc == TLFEA_OK && active > 0 && it < max_inner) {
        const int n = std::min(poll, max_inner - it);
        if (use_graph && n == poll) {
            if (!b->graph) {
                hipGraph_t g;
                TLFEA_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < poll && rc == TLFEA_OK; k++)
                    rc = nesterov_iteration(b, dt, mass_rule, u_n, v_n, f_ext, v, inner_tol, rel_tol, restart, s);
                const hipError_t ec = hipStreamEndCapture(s, &g);
                if (rc) break;
                if (ec != hipSuccess) {
                    set_error("graph capture failed: %s", hipGetErrorString(ec));
                    rc = TLFEA_EHIP;
                    break;
                }
                const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    set_error("graph instantiation failed: %s", hipGetErrorString(ei));
                    b->graph = nullptr;
                    rc = TLFEA_EHIP;
                    break;
                }
                b->graph_iters = -poll;  // not a CG graph
                b->graph_precond = -1;
            }
            TLFEA_HIP_TRY(hipGraphLaunch(b->graph, s));
        } else {
            for (int k = 0; k < n && rc == TLFEA_OK; k++)
                rc = nesterov_iteration(b, dt, mass_rule, u_n, v_n, f_ext, v, inner_tol, rel_tol, restart, s);
            if (rc) break;
        }
        it += n;
        TLFEA_HIP_TRY(hipMemcpyAsync(&active, b->d_all, sizeof(int), hipMemcpyDeviceToHost, s));
        TLFEA_HIP_TRY(hipStreamSynchronize(s));
        if (active > 0 && b->compact_ratio > 0.0) {
            TLFEA_HIP_TRY(hipMemcpy(hflags.data(), b->d_i, sizeof(int) * hflags.size(), hipMemcpyDeviceToHost));
            int64_t act_elems = 0;
            for (int i = 0; i < b->n_inst; i++) {
                done[i] = hflags[(size_t)i * NI + I_DONE];
                if (!done[i]) act_elems += b->h_elem_off[i + 1] - b->h_elem_off[i];
            }
            if ((double)act_elems < b->compact_ratio * (double)b->launched_elems)
                if ((rc = compact_lists(b, done))) break;
        }
    }
    // the graph bakes in the caller's buffers
    drop_graph(b);
    full_lists(b);
    if (rc) return rc;
    std::vector<double> S((size_t)b->n_inst * NSN);
    std::vector<int> I((size_t)b->n_inst * NI), R(b->n_inst);
    TLFEA_HIP_TRY(hipMemcpyAsync(S.data(), b->d_ns, sizeof(double) * S.size(), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(I.data(), b->d_i, sizeof(int) * I.size(), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipMemcpyAsync(R.data(), b->d_nrs, sizeof(int) * R.size(), hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    int bad = 0;
    for (int i = 0; i < b->n_inst; i++) {
        if (iters) iters[i] = I[(size_t)i * NI + I_ITERS];
        if (rel) rel[i] = S[(size_t)i * NSN + N_REL];
        if (restarts) restarts[i] = R[i];
        if (halvings) halvings[i] = I[(size_t)i * NI + I_HALVE];
        // done and not given up by the safeguard
        const int st = I[(size_t)i * NI + I_DONE] && !I[(size_t)i * NI + I_NOTPD] ? TLFEA_OK : TLFEA_ENOCONV;
        if (status) status[i] = st;
        bad += st != TLFEA_OK;
    }
    if (total_iters) *total_iters = it;
    if (bad) {
        set_error("%d of %d instances did not converge", bad, b->n_inst);
        return TLFEA_ENOCONV;
    }
    return TLFEA_OK;
}
