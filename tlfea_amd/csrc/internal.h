// Private declarations shared by the library's translation units.
#pragma once

#include "common.h"

// Node pattern + sliced block layout (built by tlfea_pattern_create, sparse.hip).
struct tlfea_pattern {
    int nn = 0, n_elem = 0, n_nodes = 0;
    int64_t nnz = 0;                  // node-level nnz
    int64_t n_slices = 0;             // 64-row slices
    int64_t sell_blocks = 0;          // padded blocks
    int* d_off = nullptr;             // node CSR offsets [n_nodes+1]
    int* d_col = nullptr;             // node CSR cols [nnz]
    int* d_slot = nullptr;            // per (e,a,b): block index in the sliced layout [n_elem*nn*nn]
    int* d_sell_row = nullptr;        // node id per slice lane [n_slices*64], -1 = padding lane
    int* d_sell_col = nullptr;        // block column node [sell_blocks]
    int64_t* d_slice_base = nullptr;  // [n_slices+1]
    int* d_slice_len = nullptr;       // [n_slices]
    int* d_row_pos = nullptr;         // node -> permuted position [n_nodes]
    unsigned char* d_mask = nullptr;  // DOF mask scratch for Dirichlet [3*n_nodes]
};

// Block-local force plan (element.hip).
struct tlfea_force_plan {
    int nn = 0, n_elem = 0, n_nodes = 0;
    int block_elems = 256;          // elements per block
    int n_blocks = 0;
    int max_u = 0;                  // max unique nodes of a block
    int64_t n_partials = 0;
    int64_t* d_blk_off = nullptr;   // [n_blocks+1] offsets into the unique-node list / partials
    int* d_unodes = nullptr;        // [n_partials] global node id of each (block, slot)
    unsigned short* d_slot = nullptr;  // AoSoA [tile][a][64]: slot of node a of element tile*64+lane
    double* d_part = nullptr;       // [n_partials*3] block partial forces
    int* d_g_off = nullptr;         // [n_nodes+1] node -> partial list
    int* d_g_idx = nullptr;         // partial positions, ascending per node (n_partials < 2^31)
    unsigned char* d_own = nullptr; // [n_partials] 1 on a node's first partial (lowest block)
    double* d_dot = nullptr;        // [2][n_blocks] per-block p.Kp and p.p shares of the fused CG
    // In-block node sums of element-major outputs (TET10 plans): a block's nodes are summed
    // in "gather positions" g, ordered by how many of the block's (element, local node)
    // pairs touch them, 64 positions (one wave) at a time.  For the wave of positions
    // [64w, 64w + 64) the table holds c_w rows of 64 u16 entries (row j, lane l: the j-th
    // pair a * 256 + e_local of position 64w + l, in ascending e_local then a; 2560 pads a
    // position with fewer than c_w pairs).
    unsigned short* d_gperm = nullptr;  // [n_partials] slot of gather position g of a block
    unsigned short* d_ginc = nullptr;   // per block: its wave tables, one after the other
    int64_t* d_ginc_off = nullptr;      // [n_blocks+1] start of a block's table (entries)
    unsigned* d_gwinfo = nullptr;       // [n_blocks][max_gw]: (table offset in the block << 8) | c_w
    int max_gw = 0;                     // ceil(max_u / 64)
    int max_ginc = 0;                   // longest block table (entries)
};

namespace tlfea_hip {

constexpr int PLAN_BLOCK = 256;  // elements per force-plan block (one thread each)

// n rounded up to a multiple of 16 (LDS carve-outs)
__host__ __device__ constexpr size_t align16(size_t n) { return (n + 15) & ~(size_t)15; }

// Planned fused force (element.hip): partials then gather into f (accumulate or overwrite).
int launch_force_planned(const tlfea_force_plan* p, const double* x, const double* y, const double* z,
                         const double* grad_t, const double* detJ_t, const tlfea_quad& quad,
                         const tlfea_material& mat, const double* v, double* f, int accumulate, const int* flags,
                         hipStream_t s);
// The same force with the gradients rebuilt from the rest positions (standard rules only,
// see is_standard_rule; LDS of 9 or, with damping, 12 arrays of max_u doubles).
bool is_standard_rule(int nn, const tlfea_quad& quad);
// elem_inst / inst_lam / inst_mu (optional): per-instance (lambda, mu) of a batch.
int launch_force_rest(const tlfea_force_plan* p, const double* x, const double* y, const double* z,
                      const double* X0, const double* Y0, const double* Z0, const tlfea_quad& quad,
                      const tlfea_material& mat, const double* v, double* f, int accumulate, const int* flags,
                      hipStream_t s, const int* elem_inst = nullptr, const double* inst_lam = nullptr,
                      const double* inst_mu = nullptr, bool x_is_disp = false);
// Options of the rest-coordinate force for implicit steps: a consistent-mass term c M dv
// added to the same block partials (c = inst_mc[instance] in a batch, mc for one body; M the
// exactly integrated mass when mass_exact, else the quadrature rule's; dv read with
// dv_stride doubles per node), the elements of instances whose inst_done[inst *
// done_stride] is set skipped, only the blocks of blk_list (n_list of them) launched, and
// the launch skipped while flags[0] | flags[1] (one body's solver flags).  SVK without
// damping only.
struct ForceMass {
    const double* dv = nullptr;
    int dv_stride = 3;
    const double* inst_mc = nullptr;
    double mc = 0.0;
    const int* flags = nullptr;
    int mass_exact = 0;
    const int* inst_done = nullptr;
    int done_stride = 0;
    const int* blk_list = nullptr;
    int n_list = 0;
    // batched Nesterov staging (nes_v non-null): per staged node n of instance i =
    // node_inst[n], y = v_k + beta_i (v_k - v_(k-1)) with beta_i = (s[0] - 1) / s[1] of the
    // instance's scalars (nes_s + i * nes_sstride), v_k = v_(k-1) = v_n when
    // nes_reset[i * done_stride] is set, y = 0 on masked DOFs; then dv = y - v_n and the
    // displacement u = u_n + dt y are formed there instead of read (the u arrays unused)
    const double* nes_v = nullptr;
    const double* nes_vkm1 = nullptr;
    const double* nes_vn = nullptr;
    const double* nes_un = nullptr;
    const unsigned char* nes_mask = nullptr;
    const int* node_inst = nullptr;
    const double* nes_s = nullptr;
    int nes_sstride = 0;
    const int* nes_reset = nullptr;
    double dt = 0.0;
};
// f_int(X + u) + c M dv as block partials in plan->d_part (no gather); u = (ux, uy, uz).
// elem_inst / inst_lam / inst_mu for a batch, else mat's (lambda, mu).
int launch_force_rest_mass_partials(const tlfea_force_plan* p, const double* ux, const double* uy, const double* uz,
                                    const double* X0, const double* Y0, const double* Z0, const tlfea_quad& quad,
                                    const tlfea_material& mat, const int* elem_inst, const double* inst_lam,
                                    const double* inst_mu, const ForceMass& fm, hipStream_t s);

// dN/dxi of every node at every quadrature point (a kernel argument).
template <typename T>
struct DnTable {
    T d[5][10][3];
    T N[5][10];  // shape function values (for the mass term)
};
DnTable<double> make_dn_table(int nn, const tlfea_quad& quad);

// Material of the tangent: one (lambda, mu) for all elements, or per instance through
// elem_inst (then inst_done[inst * done_stride], when set, skips the elements of finished
// instances).  The operator applied is k_scale * K + c M with c = mc (one body) or
// inst_mc[inst] (rho times the caller's coefficient; null = no mass term) and the
// consistent mass M of the same quadrature rule (mass_exact = 0, as the reference's
// mass_matrix_qp_kernel) or exactly integrated on straight-sided elements (mass_exact = 1).
template <typename T>
struct TanMat {
    T lam, mu;
    const int* elem_inst;
    const T* inst_lam;
    const T* inst_mu;
    const int* inst_done;
    int done_stride;
    T mc = T(0);
    const T* inst_mc = nullptr;
    T k_scale = T(1);
    int mass_exact = 0;
    int at_rest = 0;  // 1: x equals X0 (u = 0), the linear-elasticity form of the kernel
    // 1 when every element's M and det J are the same at all quadrature points
    // (tlfea_jinv_affine): the kernel then reads point 0's only
    int affine = 0;
    // 1: compressible Mooney-Rivlin (one body, FP64, no mass term) with these coefficients
    int mr = 0;
    T mu10 = T(0), mu01 = T(0), kappa = T(0);
    // Kelvin-Voigt coefficients (one body, rest positions only): the operator is K + C
    T eta = T(0), lam_d = T(0);
};

// The tangent material of one body: SVK (lambda, mu) or Mooney-Rivlin (tangent.hip).
TanMat<double> body_tanmat(const tlfea_material& m, bool at_rest, bool affine);

// Block partials of the matrix-free tangent (tangent.hip) into plan->d_part (read as T),
// for all plan blocks or, with blk_list, for the n_list blocks it names.  X0 = Y0 = Z0 =
// null: x, y, z hold the displacement itself.  With dot (all blocks, no blk_list) each
// block also writes its share of p . (K p) with masked DOFs as identity rows to
// plan->d_dot[block]: the block's unmasked DOFs' p times its partials, plus p^2 on the
// masked DOFs of the nodes whose first partial it holds (summed over blocks in order this
// is the gather's p . out); and its share of p . p (p^2 on every DOF of the nodes whose
// first partial it holds) to plan->d_dot[n_blocks + block].
template <typename T>
int launch_tangent_partials(const tlfea_force_plan* pl, const double* x, const double* y, const double* z,
                            const double* X0, const double* Y0, const double* Z0, const T* jinv_t, const T* detJ_t,
                            const tlfea_quad& quad, const TanMat<T>& mat, const unsigned char* mask, const T* p,
                            const int* flags, hipStream_t s, const int* blk_list, int n_list, bool dot = false);

// Block partials of the tangent's diagonal (FP64), gathered like the tangent's.
int launch_tangent_diag_partials(const tlfea_force_plan* pl, const double* x, const double* y, const double* z,
                                 const double* X0, const double* Y0, const double* Z0, const double* jinv_t,
                                 const double* detJ_t, const tlfea_quad& quad, const TanMat<double>& mat,
                                 hipStream_t s);

// Matrix-free tangent action (FP64, one body) at positions x (rest X0, inverse Jacobians
// jinv_t); dot_part (optional) receives per-block partials of p . out (blocks of 256 nodes).
// at_rest: the caller guarantees x == X0 (the linear-elasticity form of the kernel).
int launch_tangent_planned(const tlfea_force_plan* pl, const double* x, const double* y, const double* z,
                           const double* X0, const double* Y0, const double* Z0, const double* jinv_t,
                           const double* detJ_t, const tlfea_quad& quad, const tlfea_material& mat,
                           const unsigned char* mask, const double* p, double* out, double* dot_part,
                           const int* flags, hipStream_t s, bool at_rest = false, bool affine = false);
// true when x, y, z equal X0, Y0, Z0 on every node (same arrays, or equal values: one
// device pass over d_scratch, one device int, and a synchronisation of s).
int positions_at_rest(int n_nodes, const double* x, const double* y, const double* z, const double* X0,
                      const double* Y0, const double* Z0, int* d_scratch, hipStream_t s, bool* at_rest);

// Element precompute (element.hip): any of grad_t / detJ_t / jinv_t may be null.
int launch_precompute(int nn, int n_elem, const int* conn, const double* x, const double* y, const double* z,
                      const tlfea_quad& quad, double* grad_t, double* detJ_t, double* jinv_t, hipStream_t s);

// Fused FP64 force launch (element.hip).  With `flags` non-null the kernel returns at once
// when flags[0] | flags[1] is set (used inside the iterative solvers' queued iterations).
int launch_force(int nn, int n_elem, const int* conn, const double* x, const double* y, const double* z,
                 const double* grad_t, const double* detJ_t, const tlfea_quad& quad, const tlfea_material& mat,
                 const double* v, double* f, const int* flags, hipStream_t s);

// Sum of one node's block partials part[3 idx[k] + d] over k in [k0, k1), in ascending k.
// The slot indices of four entries are loaded before any of their partials, so a node
// costs two memory round trips per four entries rather than two per entry; the additions
// run in the same order as a plain loop, so the sums are the same to the bit.
template <typename T>
__device__ __forceinline__ void gather_slots3(const int* __restrict__ idx, const T* __restrict__ part, int k0,
                                              int k1, T& s0, T& s1, T& s2) {
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
        int64_t j[4];
#pragma unroll
        for (int u = 0; u < 4; u++) j[u] = 3 * (int64_t)idx[k + u];
        T a[4][3];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a[u][0] = part[j[u]];
            a[u][1] = part[j[u] + 1];
            a[u][2] = part[j[u] + 2];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            s0 += a[u][0];
            s1 += a[u][1];
            s2 += a[u][2];
        }
    }
    for (; k < k1; k++) {
        const T* pp = part + 3 * (int64_t)idx[k];
        s0 += pp[0];
        s1 += pp[1];
        s2 += pp[2];
    }
}

}  // namespace tlfea_hip
