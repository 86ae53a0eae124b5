/*
 * tlfea_hip.h -- C-ABI of the MI355X (gfx950) TLFEA solve-path library (libtlfea_hip.so).
 *
 * Plain C: POD structs, device pointers, sizes, hipStream_t (passed as void*), int status.
 * The C++ drop-in classes (include/tlfea/) are the only intended callers; tests and the
 * benchmark also bind it through ctypes.  Nothing here allocates inside a hot call: the
 * calls that need persistent device state (patterns, CG workspaces) return opaque handles
 * from setup calls.
 *
 * Device data layouts (all FP64 unless noted):
 *   conn      int32, column-major conn[e + a*n_elem]         (the reference's MatrixXi Map)
 *   x, y, z   SoA node positions
 *   f, v, u   AoS per DOF [3*node + d]
 *   grad_t    AoSoA tiles of 64 elements: grad_t[((tile*nq + q)*nn*3 + a + nn*d)*64 + lane],
 *             element e = tile*64 + lane, entry = dN_a/dX_d at quadrature point q
 *   detJ_t    detJ_t[(tile*nq + q)*64 + lane]
 *   qp tensors (F, P, Fdot, P_vis): qp_t[((tile*nq + q)*9 + i + 3*j)*64 + lane]
 * Padded lanes (e >= n_elem) hold zeros.  tlfea_padded_elems() gives the padded count.
 */
#ifndef TLFEA_HIP_H
#define TLFEA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TLFEA_OK 0
#define TLFEA_EINVAL (-1)    /* bad argument (size, null pointer, unsupported element) */
#define TLFEA_EHIP (-2)      /* HIP runtime error */
#define TLFEA_ENOTPD (-3)    /* CG met p.Kp <= 0 (LinearStaticSolver.cu:404-407)      */
#define TLFEA_ECAPACITY (-4) /* a setup limit (row length, index range) was exceeded */
#define TLFEA_ENOCONV (-5)   /* iteration limit reached without convergence          */

#define TLFEA_TILE 64 /* elements per AoSoA tile = one wavefront */

/* Quadrature rule, passed by value into kernels.  n = 5 (TET10 Keast) or 1 (TET4). */
typedef struct tlfea_quad {
    int n;
    double x[5], y[5], z[5], w[5];
} tlfea_quad;

/* Material: SVK (lambda, mu) or, with model = 1, compressible Mooney-Rivlin (mu10, mu01,
 * kappa; src/materials/MooneyRivlin.cuh, selected as material_model == 1 in compute_p,
 * FEAT10DataFunc.cuh:274-275), plus Kelvin-Voigt damping (eta, lambda_d).  Zero-filled
 * trailing fields give SVK. */
typedef struct tlfea_material {
    double lambda, mu, eta_damp, lambda_damp;
    int model; /* 0 SVK, 1 Mooney-Rivlin */
    double mu10, mu01, kappa;
} tlfea_material;

/* ---- general ---------------------------------------------------------------------- */
const char* tlfea_version(void);
const char* tlfea_last_error(void);
int tlfea_device_synchronize(void);
int64_t tlfea_padded_elems(int n_elem);
/* Replaces the reference's Quadrature::tet5pt_* / tet1pt_* tables
 * (src/utils/quadrature_utils.h:125-179). */
int tlfea_quadrature(int nn, tlfea_quad* out);

/* ---- element precompute (setup) ---------------------------------------------------- */
/* Replaces dn_du_pre_kernel / GPU_FEAT10_Data::CalcDnDuPre (src/elements/FEAT10Data.cu:92-198,
 * 277-285) and dn_du_pre_feat4_kernel (src/elements/FEAT4Data.cu:98-190).  nn = 10 or 4. */
int tlfea_precompute_grad(int nn, int n_elem, const int* conn, const double* x, const double* y, const double* z,
                          const tlfea_quad* quad, double* grad_t, double* detJ_t, void* stream);
/* Inverse-transposed Jacobians M = J^-T per (element, q) from the same J and pivoting, so
 * that g_a = M dN_a/dxi (a pivot below 1e-14 gives M = 0).  Layout
 * jinv_t[((tile*nq + q)*9 + 3i + k)*64 + lane] = M[i][k], padded_elems(n_elem)*nq*9
 * doubles.  Not in the reference: the compact input of the matrix-free tangent. */
int tlfea_precompute_jinv(int nn, int n_elem, const int* conn, const double* x, const double* y, const double* z,
                          const tlfea_quad* quad, double* jinv_t, void* stream);
/* *affine = 1 when every element has the same M and det J at all quadrature points to
 * 1e-12 relative (straight-sided elements: J is constant, the per-point values differ by
 * the rounding of sum_a X_a (x) dN_a/dxi only).  Setup call (synchronises the stream).  With
 * affine = 1 the matrix-free tangent reads point 0's M and det J only.  Not in the
 * reference. */
int tlfea_jinv_affine(int nn, int n_elem, const double* jinv_t, const double* detJ_t, int* affine, void* stream);

/* ---- stress and internal force ---------------------------------------------------- */
/* Replaces calc_p_kernel / compute_p (src/elements/FEAT10Data.cu:287-305,
 * src/elements/FEAT10DataFunc.cuh:86-291; TET4 src/elements/FEAT4DataFunc.cuh:33-200):
 * stores F, P = P_svk + P_vis, Fdot, P_vis at every quadrature point (any output may be
 * NULL).  v == NULL (or zero damping) stores Fdot = P_vis = 0. */
int tlfea_calc_p(int nn, int n_elem, const int* conn, const double* x, const double* y, const double* z,
                 const double* grad_t, const tlfea_material* mat, const double* v, double* F_t, double* P_t,
                 double* Fdot_t, double* Pvis_t, void* stream);

/* Replaces compute_internal_force_kernel / CalcInternalForce
 * (src/elements/FEAT10Data.cu:307-323, src/elements/FEAT10DataFunc.cuh:404-460):
 * f[3*node+d] += sum_q P_q g_a(q) detJ_q w_q.  Accumulates onto f, as the reference does. */
int tlfea_fint_from_p(int nn, int n_elem, const int* conn, const double* P_t, const double* grad_t,
                      const double* detJ_t, const tlfea_quad* quad, double* f, void* stream);

/* Fused P + f_int (the solver loops' force evaluation, src/solvers/SyncedNesterov.cu:180-207):
 * computes the SVK (+ Kelvin-Voigt when v != NULL) stress per quadrature point in registers
 * and scatters the element nodal forces; nothing per quadrature point is stored.
 * accumulate == 0 zeroes f first. */
int tlfea_internal_force(int nn, int n_elem, int n_nodes, const int* conn, const double* x, const double* y,
                         const double* z, const double* grad_t, const double* detJ_t, const tlfea_quad* quad,
                         const tlfea_material* mat, const double* v, double* f, int accumulate, void* stream);

/* ---- block-local force plan (setup; opaque handle owns device memory) ------------ */
/* Elements are cut into blocks of up to 256 consecutive elements; each block gets the
 * sorted list of the nodes it touches and, per (element, local node), the slot of that
 * node in the list.  The planned force kernel stages the block's node positions (and
 * velocities) in LDS, accumulates the element nodal forces in LDS and writes one partial
 * per (block, node) with plain stores; a node-gather kernel then sums each node's
 * partials in a fixed order.  No global atomics; the in-block LDS accumulation order
 * follows wave scheduling, so repeated calls agree to rounding, not bit for bit. */
typedef struct tlfea_force_plan tlfea_force_plan;
int tlfea_force_plan_create(int nn, int n_elem, int n_nodes, const int* conn, tlfea_force_plan** out, void* stream);
int tlfea_force_plan_destroy(tlfea_force_plan* p);
/* elements per block, blocks, total partials (sum over blocks of unique nodes), max
 * unique nodes in one block */
int tlfea_force_plan_sizes(const tlfea_force_plan* p, int* block_elems, int* n_blocks, int64_t* n_partials,
                           int* max_block_nodes);
/* Same result as tlfea_internal_force (SVK + Kelvin-Voigt when v != NULL). */
int tlfea_internal_force_planned(const tlfea_force_plan* plan, const double* x, const double* y, const double* z,
                                 const double* grad_t, const double* detJ_t, const tlfea_quad* quad,
                                 const tlfea_material* mat, const double* v, double* f, int accumulate, void* stream);
/* CalcInternalForce (src/elements/FEAT10Data.cu:307-323, FEAT4Data.cu) with the shape-function
 * gradients of CalcDnDuPre (FEAT10Data.cu:92-198) rebuilt per quadrature point from the
 * rest positions X0, Y0, Z0 instead of read from tlfea_precompute_grad's arrays: the same
 * f_int as tlfea_internal_force_planned (the strain from the displacement gradient x - X,
 * so at small strains it is the more accurate of the two).  Standard rules only (TET10:
 * 5 points, TET4: 1 point; TLFEA_EINVAL otherwise); TLFEA_ECAPACITY when a plan block
 * touches more nodes than 9 (12 with damping) LDS arrays of doubles hold. */
int tlfea_internal_force_rest(const tlfea_force_plan* plan, const double* x, const double* y, const double* z,
                              const double* X0, const double* Y0, const double* Z0, const tlfea_quad* quad,
                              const tlfea_material* mat, const double* v, double* f, int accumulate, void* stream);

/* Matrix-free tangent action y = K(x) p, the operator that compute_hessian_assemble_csr +
 * SolveLinearSystemCG's SpMV apply (src/elements/FEAT10DataFunc.cuh:512-768,
 * src/materials/SVK.cuh:41-67, src/solvers/LinearStaticSolver.cu:360-371) without
 * assembling it: per quadrature point F = I + sum_a u_a (x) g_a (u = x - X0),
 * dF = sum_a p_a (x) g_a, dE = (dF^T F + F^T dF)/2, dP = dF S + F (lambda tr(dE) I + 2 mu dE),
 * y_a = sum_q dP g_a dV, with g_a rebuilt from jinv_t (tlfea_precompute_jinv).  Equal to
 * the assembled K(x) p up to rounding; reads ~420 B per TET10 element (M, detJ, slots)
 * instead of ~3 kB of assembled 3x3 blocks.  With a DOF mask (uint8 [3*n_nodes],
 * 1 = fixed) the masked rows return y = p (the Dirichlet identity rows of
 * tlfea_apply_dirichlet) and the masked columns are ignored (p read as 0); NULL = no
 * mask.  With Kelvin-Voigt damping in mat the operator is the reference's K + C
 * (FEAT10DataFunc.cuh:686-767): C p = F (lambda_d tr(dE) I + 2 eta dE) per point, dE as
 * above, added to dS; at F = I that makes K + C the linear-elasticity operator with
 * lambda + lambda_d and mu + eta (Mooney-Rivlin: the same C p added to A : dF).
 * affine (as reported by tlfea_jinv_affine; 0 is always valid): M and det J of quadrature
 * point 0 only are read.  When x, y, z are the arrays X0, Y0, Z0 themselves the kernel
 * runs its linear-elasticity form (u = 0: H = S = 0, dP = dS), the same values without
 * the u terms. */
int tlfea_tangent_apply_planned(const tlfea_force_plan* plan, const double* x, const double* y, const double* z,
                                const double* X0, const double* Y0, const double* Z0, const double* jinv_t,
                                const double* detJ_t, int affine, const tlfea_quad* quad, const tlfea_material* mat,
                                const unsigned char* fixed_dof_mask, const double* p, double* out, void* stream);

/* FP32 variant of the fused force kernel, displacement form F = I + sum_a u_a (x) g_a:
 * u = x - X is passed as three float arrays, grad32_t / dV32_t are float copies of grad_t
 * and detJ_t*w (tlfea_grad_to_f32).  f is float. */
int tlfea_grad_to_f32(int nn, int n_elem, const double* grad_t, const double* detJ_t, const tlfea_quad* quad,
                      float* grad32_t, float* dV32_t, void* stream);
int tlfea_internal_force_f32(int nn, int n_elem, int n_nodes, const int* conn, const float* ux, const float* uy,
                             const float* uz, const float* grad32_t, const float* dV32_t, const tlfea_material* mat,
                             float* f, int accumulate, void* stream);

/* Layout converters for the drop-in Retrieve*ToCPU calls: AoSoA -> the reference's layouts
 * grad[(e*nq+q)*nn*3 + a + nn*d], detJ[e*nq+q], qp[(e*nq+q)*9 + i + 3j]. */
int tlfea_grad_to_ref(int nn, int n_elem, const double* grad_t, const double* detJ_t, double* grad_ref,
                      double* detJ_ref, void* stream);
int tlfea_qp_to_ref(int nn, int n_elem, const double* qp_t, double* qp_ref, void* stream);

/* ---- sparsity patterns (setup; opaque handle owns device memory) ------------------ */
typedef struct tlfea_pattern tlfea_pattern;
/* Builds, on the GPU from element->node incidence (no key sort), the node-level pattern
 * of src/elements/FEAT10Data.cu:34-74,361-417 and the DOF-level pattern of
 * src/solvers/LinearStaticSolver.cu:75-122,240-294 (= node pattern expanded 3x3), plus the
 * element -> slot maps used for scatter-free-of-search assembly and the sliced (64-row)
 * block layout used by the SpMV. */
int tlfea_pattern_create(int nn, int n_elem, int n_nodes, const int* conn, tlfea_pattern** out, void* stream);
/* As tlfea_pattern_create with an explicit row-sort window (rows are sorted by length
 * inside consecutive windows of `window` rows before slicing; 1 keeps the mesh order). */
int tlfea_pattern_create_ex(int nn, int n_elem, int n_nodes, const int* conn, int window, tlfea_pattern** out,
                            void* stream);
int tlfea_pattern_destroy(tlfea_pattern* p);
/* node_nnz = mass-matrix nnz, dof_nnz = 9*node_nnz = stiffness nnz, sell_blocks = padded
 * block count of the sliced layout (values per matrix: 9*sell_blocks for K, sell_blocks
 * for M). */
int tlfea_pattern_sizes(const tlfea_pattern* p, int64_t* node_nnz, int64_t* dof_nnz, int64_t* sell_blocks);
/* Exact reference CSR patterns (device outputs, int32). */
int tlfea_pattern_node_csr(const tlfea_pattern* p, int* offsets, int* cols, void* stream);
int tlfea_pattern_dof_csr(const tlfea_pattern* p, int* offsets, int* cols, void* stream);

/* ---- mass and stiffness assembly ------------------------------------------------- */
/* Replaces mass_matrix_qp_kernel / CalcMassMatrix (src/elements/FEAT10Data.cu:200-271,
 * 341-359; TET4 src/elements/FEAT4Data.cu:197-243).  Writes the sliced-layout values
 * m_sell[sell_blocks] (zeroed first). */
int tlfea_mass_assemble(const tlfea_pattern* p, int nn, int n_elem, const int* conn, const double* detJ_t,
                        const tlfea_quad* quad, double rho, double* m_sell, void* stream);
/* Sliced-layout node matrix -> values in the node CSR order of tlfea_pattern_node_csr. */
int tlfea_mass_to_csr(const tlfea_pattern* p, const double* m_sell, double* vals_csr, void* stream);
/* Replaces assemble_stiffness_kernel / compute_hessian_assemble_csr
 * (src/solvers/LinearStaticSolver.cu:128-141,300-310; src/elements/FEAT10DataFunc.cuh:512-768;
 * src/materials/SVK.cuh:41-67): K = h * K_svk(x) + C_kv(x), written into the sliced 3x3
 * block layout k_sell[9*sell_blocks] (zeroed first). */
int tlfea_stiffness_assemble(const tlfea_pattern* p, int nn, int n_elem, const int* conn, const double* x,
                             const double* y, const double* z, const double* grad_t, const double* detJ_t,
                             const tlfea_quad* quad, const tlfea_material* mat, double h, double* k_sell,
                             void* stream);
/* Sliced block matrix -> values in the DOF CSR order of tlfea_pattern_dof_csr. */
int tlfea_stiffness_to_csr(const tlfea_pattern* p, const double* k_sell, double* vals_csr, void* stream);
/* Replaces mark_fixed_dofs_kernel + apply_bcs_kernel (src/solvers/LinearStaticSolver.cu:
 * 147-184,316-342): fixed DOF rows become identity rows with f = 0, free rows get the fixed
 * columns zeroed.  fixed_nodes: device int32[n_fixed]. */
int tlfea_apply_dirichlet(const tlfea_pattern* p, int n_fixed, const int* fixed_nodes, double* k_sell, double* f,
                          void* stream);

/* ---- sparse / dense kernels -------------------------------------------------------- */
/* y = K x on the sliced 3x3 block layout (replaces cusparseSpMV, LinearStaticSolver.cu:398). */
int tlfea_spmv_block(const tlfea_pattern* p, const double* k_sell, const double* x, double* y, void* stream);
/* FP32 y = K32 x on the sliced layout (the mixed-precision CG's operator). */
int tlfea_spmv_block_f32(const tlfea_pattern* p, const float* k32, const float* x, float* y, void* stream);
/* y = M x on the node matrix applied per component (x, y AoS 3 per node). */
int tlfea_spmv_node3(const tlfea_pattern* p, const double* m_sell, const double* x, double* y, void* stream);
/* Plain scalar CSR SpMV (the reference's matrix format), y = A x. */
int tlfea_spmv_csr(int n_rows, const int* offsets, const int* cols, const double* vals, const double* x, double* y,
                   void* stream);
/* BLAS-1 with device-resident results (replaces cublasDdot/Daxpy/Dscal/Dnrm2,
 * LinearStaticSolver.cu:379-428, SyncedAdamWNocoop.cu:34).  Reductions are two-level and
 * deterministic.  `scratch` must hold tlfea_reduce_scratch_size(n) doubles. */
int64_t tlfea_reduce_scratch_size(int64_t n);
int tlfea_dot(int64_t n, const double* a, const double* b, double* out, double* scratch, void* stream);
int tlfea_nrm2(int64_t n, const double* a, double* out, double* scratch, void* stream);
int tlfea_axpy(int64_t n, double alpha, const double* x, double* y, void* stream);
int tlfea_scal(int64_t n, double alpha, double* x, void* stream);

/* ---- CG (replaces LinearStaticSolver::SolveLinearSystemCG, LinearStaticSolver.cu:348-443) */
typedef struct tlfea_cg tlfea_cg;
int tlfea_cg_create(const tlfea_pattern* p, tlfea_cg** out);
int tlfea_cg_destroy(tlfea_cg* c);
/* Unpreconditioned CG from u = 0, stop when sqrt(rr/rr0) < tol (the reference's rule),
 * TLFEA_ENOTPD on p.Kp <= 0.  Scalars stay on the device; the host reads the convergence
 * flag every `check_every` iterations only. */
int tlfea_cg_solve(tlfea_cg* c, const double* k_sell, const double* f, double* u, double tol, int max_iter,
                   int check_every, int* iters_out, double* rel_res_out, void* stream);

/* ---- constraints (FEAT10DataFunc.cuh:470-481, FEAT10Data.cu:325-339) -------------- */
/* c[3k+d] = x_d(fixed[k]) - x_ref,d(fixed[k]) for k < n_fixed. */
int tlfea_constraint_eval(int n_fixed, const int* fixed_nodes, const double* x, const double* y, const double* z,
                          const double* xref, const double* yref, const double* zref, double* c, void* stream);

/* ---- matrix-free CG (not in the reference) ----------------------------------------- */
/* The reference's CG (same sqrt(rr/rr0) < tol rule, TLFEA_ENOTPD on p.Kp <= 0) with the
 * operator K(x) applied by tlfea_tangent_apply_planned instead of an assembled matrix:
 * the linear-static system of LinearStaticSolver::Solve after Dirichlet elimination.
 * When x equals X0 on every node (checked once per call) the operator is applied in its
 * linear-elasticity form.  With Kelvin-Voigt damping the operator is K + C. */
int tlfea_cg_solve_matfree(tlfea_cg* c, const tlfea_force_plan* plan, const double* x, const double* y,
                           const double* z, const double* X0, const double* Y0, const double* Z0,
                           const double* jinv_t, const double* detJ_t, int affine, const tlfea_quad* quad,
                           const tlfea_material* mat, const unsigned char* fixed_dof_mask, const double* f, double* u,
                           double tol, int max_iter, int check_every, int* iters_out, double* rel_res_out,
                           void* stream);
/* No counterpart in the reference.  tlfea_cg_solve_matfree without the tlfea_cg handle,
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same kernels, arguments, stopping
 * rule and status codes, with the CG's own workspace (r, p, K p, reduction partials,
 * scalars, flags) in a static device array of the library, 72 MiB per GPU, held until the
 * process ends.  It calls no hipMalloc or hipFree.  Bodies of more than
 * TLFEA_NOALLOC_MAX_NODES nodes return TLFEA_ECAPACITY.  The force plan (a setup product of
 * tlfea_force_plan_create) still carries the operator's block partials.  Every call on a
 * GPU shares that GPU's array and the plan's partials: calls that overlap on one GPU (two
 * streams, two host threads) race. */
#define TLFEA_NOALLOC_MAX_NODES 1048576
int tlfea_cg_solve_matfree_noalloc(const tlfea_force_plan* plan, const double* x, const double* y, const double* z,
                                   const double* X0, const double* Y0, const double* Z0, const double* jinv_t,
                                   const double* detJ_t, int affine, const tlfea_quad* quad, const tlfea_material* mat,
                                   const unsigned char* fixed_dof_mask, const double* f, double* u, double tol,
                                   int max_iter, int check_every, int* iters_out, double* rel_res_out, void* stream);
/* DOF mask (uint8 [3*n_nodes], 1 = DOF of a fixed node) from a fixed-node list. */
int tlfea_fixed_dof_mask(int n_nodes, int n_fixed, const int* fixed_nodes, unsigned char* mask, void* stream);

/* ---- batches of independent meshes (not in the reference) ---------------------------- */
/* Instances are concatenated: instance i owns elements elem_off[i]..elem_off[i+1]-1 and
 * nodes node_off[i]..node_off[i+1]-1 (host arrays of n_inst+1, starting at 0); conn
 * (device, column-major [nn][n_elem]) holds concatenated node ids; x/y/z are the rest
 * positions; lambda/mu (host, per instance) the Lame constants; fixed_dof_mask (device,
 * uint8 [3*n_nodes]) marks fixed DOFs.  Inputs are copied; the handle owns its memory. */
typedef struct tlfea_batch tlfea_batch;
int tlfea_batch_create(int nn, int n_inst, const int* elem_off, const int* node_off, const int* conn,
                       const double* x, const double* y, const double* z, const double* lambda, const double* mu,
                       const unsigned char* fixed_dof_mask, tlfea_batch** out, void* stream);
int tlfea_batch_destroy(tlfea_batch* b);
int tlfea_batch_sizes(const tlfea_batch* b, int* n_inst, int* n_elem, int* n_nodes, int* n_chunks);
/* Nodes per reduction chunk (>= 64): the granularity of the per-instance dot products. */
int tlfea_batch_set_chunk(tlfea_batch* b, int chunk_nodes);
/* Autotuner (not in the reference): times batched CG iterations for every chunk size in
 * {256, 512, 1024, 2048, 4096} nodes and graph-replay length in {16, 32, 64} iterations
 * (runs of n and 2n iterations from u = 0 on load f, n >= trial_iters, no instance
 * finishing, each timed twice with the faster kept; the difference cancels the per-solve
 * graph capture, and where noise makes it non-positive the 2n run's mean per iteration
 * stands in), keeps the fastest pair
 * (chunk via tlfea_batch_set_chunk, replay length as the default of solves that pass
 * poll <= 0), and with full_solves != 0 then times full solves to tol for compaction
 * ratios {0.35, 0.5, 0.65, 0.8, 0.9} and keeps the fastest.  u is a scratch output
 * [3*n_nodes]; the batch's solution state is overwritten. */
typedef struct tlfea_batch_tuning {
    int chunk_nodes;    /* nodes per vector-kernel chunk */
    int poll;           /* iterations per graph replay between host checks */
    double compaction;  /* compaction ratio */
    double us_per_iter; /* time of one batched iteration, all instances active */
    double solve_s;     /* full-solve time at the chosen ratio (0 without full_solves) */
} tlfea_batch_tuning;
int tlfea_batch_autotune(tlfea_batch* b, const double* f, double* u, int trial_iters, double tol, int full_solves,
                         tlfea_batch_tuning* out, void* stream);
/* Finished instances are dropped from the launches once the unfinished ones hold less
 * than `ratio` of the launched elements (default 0.6; 0 disables). */
int tlfea_batch_set_compaction(tlfea_batch* b, double ratio);
/* out = K p for every instance at its rest configuration, Dirichlet rows/columns masked
 * (the operator of tlfea_tangent_apply_planned per instance). */
int tlfea_batch_tangent_apply(tlfea_batch* b, const double* p, double* out, void* stream);
/* The batch's operator (K, plus c M once tlfea_batch_set_mass has added it) in FP32
 * (inverse Jacobians, det J, Lame parameters and c rounded to FP32, accumulation in FP64
 * LDS per block): the operator of the FP32 preconditioner stages. */
int tlfea_batch_tangent_apply_f32(tlfea_batch* b, const float* p, float* out, void* stream);
/* Adds the consistent mass M, applied without assembly, to the operator of
 * tlfea_batch_cg_solve: it becomes K + c M per instance, with rho_coef[i] = rho_i * c
 * (host [n_inst]); NULL removes it.  Used by implicit time steps (c = 1/dt^2).
 * rule 0: the quadrature rule of the mass assembly (mass_matrix_qp_kernel,
 * FEAT10Data.cu:200-271; FEAT4Data.cu:197-243), whose TET10 mass is indefinite (negative
 * centroid weight) and TET4 mass rank one, so K + c M loses definiteness once c M
 * dominates K; rule 1: the exactly integrated mass of straight-sided elements (symmetric
 * positive definite, same row sums).  Removes the preconditioner (tlfea_batch_precond_setup
 * and tlfea_batch_coarse_setup were for the previous operator: set them up again). */
int tlfea_batch_set_mass(tlfea_batch* b, const double* rho_coef, int rule);
/* out = (k_scale K + with_mass * c M) p for every instance; masked != 0 applies the
 * Dirichlet elimination (fixed rows return p, fixed columns are dropped), masked = 0 the
 * full operator (e.g. the consistent gravity load M g, which sums over all nodes). */
int tlfea_batch_apply(tlfea_batch* b, double k_scale, int with_mass, int masked, const double* p, double* out,
                      void* stream);
/* Displacement of every instance from its rest positions (device SoA [n_nodes] each,
 * copied): from then on the batch's operator (tlfea_batch_cg_solve, tlfea_batch_apply, the
 * preconditioner's diagonal) is the tangent K(X + u) (compute_hessian_assemble_csr,
 * FEAT10DataFunc.cuh:512-768, per instance) instead of K at rest; NULLs return to rest.
 * The kernels take u itself, not x = X + u: a displacement many orders below the
 * coordinates would lose its digits in x - X.  For Newton iterations of implicit steps. */
int tlfea_batch_set_displacement(tlfea_batch* b, const double* ux, const double* uy, const double* uz, void* stream);
/* f = f_int at the current positions (CalcInternalForce, FEAT10Data.cu:307-323, with each
 * instance's material; gradients rebuilt from the rest positions as in
 * tlfea_internal_force_rest), all DOFs (no Dirichlet masking), overwritten. */
int tlfea_batch_internal_force(tlfea_batch* b, double* f, void* stream);
/* out[i] = a . c over instance i's DOFs (device double[n_inst]; all DOFs, no masking), by
 * chunk partials summed per instance in chunk order (deterministic).  Per-instance norms
 * for the Newton iterations of implicit steps.  Not in the reference. */
int tlfea_batch_dot(tlfea_batch* b, const double* a, const double* c, double* out, void* stream);
/* FP32 Jacobi-Chebyshev preconditioner for tlfea_batch_cg_solve(precond = 1) of the batch's
 * operator A (K, or K + c M with a mass term, at the current positions): the diagonal D of
 * A (FP64, matrix-free), lambda_max(D^-1 A) per instance by `power_iters`
 * FP32 power iterations (scaled by `safety` >= 1), and `degree` (at most 64) Chebyshev steps
 * on [lambda_max / ratio, lambda_max] applied with the FP32 operator.  degree = 0 removes it.
 * lambda_max (host [n_inst], may be NULL) receives the unscaled estimates. */
int tlfea_batch_precond_setup(tlfea_batch* b, int degree, double ratio, int power_iters, double safety,
                              double* lambda_max, void* stream);
/* Two-level coarse correction for the FP32 stage, not in the reference: the nodes of each
 * instance are grouped into planes across its longest axis (the mesh's lattice planes, or
 * for other meshes equal-width slabs half as thick as the longest element along that axis,
 * so that no element spans more than three); the coarse space holds the 6 rigid motions of
 * each plane's free nodes and, on request where W <= 2 (tlfea_batch_set_coarse_modes), five
 * in-plane fields of the plane's cross-section coordinates (s, t): the quadratic warping
 * (s t, (t^2 - s^2)/2) and ((s^2 - t^2)/2, s t) that bending about either axis brings with
 * the Poisson effect, the uniform stretch (s, t) and the two shears (t, s), (s, -t).
 * A_c = P^T A P (A = K, or K + c M with a mass term) is assembled by probing with the FP64
 * operator
 * and factored per instance as a block-tridiagonal matrix (blocks of W planes, W = the
 * widest plane span of an element, at most 3: TLFEA_ECAPACITY beyond).  tlfea_batch_cg_solve
 * (precond = 1) then applies the additive two-level preconditioner z = P A_c^-1 P^T r +
 * Chebyshev(r) inside the flexible FP64 CG: the beams' bending modes, which the Chebyshev
 * stage cannot reach, go to the coarse solve.  Needs tlfea_batch_precond_setup first;
 * enable = 0 removes it.  n_coarse (may be NULL) receives the coarse dimension. */
int tlfea_batch_coarse_setup(tlfea_batch* b, int enable, int* n_coarse, void* stream);
/* Coarse modes per plane for the next tlfea_batch_coarse_setup: 6 (the default: the rigid
 * motions) or 11 (the rigid motions and the five in-plane fields, used where elements span
 * at most 2 planes, else the rigid motions alone).  Not in the reference. */
int tlfea_batch_set_coarse_modes(tlfea_batch* b, int modes);
/* z = P A_c^-1 P^T r for every instance (FP64; r and z [3 n_nodes] on the device): the
 * coarse part of the preconditioner on its own. */
int tlfea_batch_coarse_apply(tlfea_batch* b, const double* r, double* z, void* stream);
/* u += Chebyshev(f - K u), `sweeps` times: the FP32 stage of tlfea_batch_precond_setup used
 * as a smoother of an initial guess (it damps the error components of the upper part of
 * the spectrum of D^-1 K). */
int tlfea_batch_smooth(tlfea_batch* b, const double* f, double* u, int sweeps, void* stream);
/* Linear-static CG for every instance at once: K_i u_i = f_i (fixed DOFs eliminated),
 * the reference's CG per instance with its own alpha/beta, stopping when
 * ||f_i - K_i u_i|| / ||f_i|| < tol on the free DOFs.  use_guess != 0 starts from u
 * (else from 0).  Finished instances stop updating; the host polls every `poll`
 * iterations (poll <= 0: the batch's default, 32 or the autotuner's choice), which are
 * replayed from a captured hipGraph when use_graph != 0.  With
 * precond = 1 the CG is the flexible (Polak-Ribiere) variant preconditioned by the FP32
 * Chebyshev stage of tlfea_batch_precond_setup; u, r, alpha, beta and the stopping test
 * stay FP64.
 * Per-instance results go to the host arrays iters / rel_res / status (TLFEA_OK,
 * TLFEA_ENOCONV, TLFEA_ENOTPD); total_iters = iterations run for the batch.  Returns
 * TLFEA_ENOCONV if any instance failed. */
int tlfea_batch_cg_solve(tlfea_batch* b, const double* f, double* u, int use_guess, double tol, int max_iter,
                         int poll, int use_graph, int precond, int* iters, double* rel_res, int* status,
                         int* total_iters, void* stream);
/* Batched SyncedNesterov: one implicit step for every instance at once, replacing the
 * inner loop of SyncedNesterovSolver::Solve (src/solvers/SyncedNesterov.cu:97-360, the
 * gradient of :32-95) run per body.  Each instance minimises over v the step equation
 *   g(v) = M (v - v_n) / dt + f_int(X + u_n + dt v) - f_ext = 0
 * with its own step size alpha[i] (host [n_inst]), momentum sequence and stopping test:
 * y = v_k + beta (v_k - v_(k-1)), v_next = y - alpha g(y), beta = (t_k - 1) / t_(k+1).
 * M is the consistent mass with density rho[i] (host [n_inst]; mass_rule 0 the 5-point
 * rule of mass_matrix_qp_kernel, 1 exactly integrated), f_int the SVK force of each
 * instance's material from the rest positions (as tlfea_batch_internal_force), u_n, v_n,
 * f_ext device AoS [3 n_nodes] (the displacement and velocity at the step start, the
 * load).  Fixed DOFs are eliminated (v = 0 there) instead of being constrained by the
 * augmented-Lagrangian outer loop.  v (device, [3 n_nodes]) starts from itself
 * (use_guess != 0) or from v_n and returns each instance's last iterate.  Stopping tests:
 * the reference's, | |g| - |g_prev| | < inner_tol or | |v_next| - |v_k| | < inner_tol after
 * the first iteration (inner_tol > 0), and the relative |g| <= rel_tol |g_0| (rel_tol > 0;
 * g_0 the gradient at the first iterate); restart != 0 restarts the momentum when
 * g.(v_next - v_k) > 0.  Iterations are replayed from a captured hipGraph of `poll` of them
 * (use_graph != 0), finished instances drop out as in tlfea_batch_cg_solve.  Safeguard
 * (not in the reference): an instance whose |g| turns non-finite or exceeds 1e4 |g_0| halves
 * its step size and starts again from v_n (at most 40 times).  Per instance (host arrays,
 * may be NULL): iterations, |g| / |g_0|, restarts, step halvings, status (TLFEA_ENOCONV
 * when max_inner or the halvings ran out).  Standard quadrature rules only. */
int tlfea_batch_nesterov_solve(tlfea_batch* b, double dt, const double* rho, int mass_rule, const double* alpha,
                               const double* u_n, const double* v_n, const double* f_ext, double* v, int use_guess,
                               double inner_tol, double rel_tol, int max_inner, int restart, int poll, int use_graph,
                               int* iters, double* rel, int* restarts, int* halvings, int* status, int* total_iters,
                               void* stream);

/* ---- learned warm start (not in the reference) -------------------------------------- */
/* A message-passing network on the mesh node graph (CSR off/col of tlfea_pattern_node_csr,
 * diagonal included): h0 = relu(x W_in^T + b_in); per layer m = mean_{j in N(i)} h_j,
 * h = h + relu([h, m] W_l^T + b_l); out = h W_out^T + b_out.  Width 64, 16 input features,
 * 3 outputs (padded to 16).  BF16 weights and activations on the matrix cores with FP32
 * accumulation; FP32 master weights and Adam.  Parameters (FP32, n_params) are packed as
 * W_in [64x16], b_in [64], per layer W_l [64x128], b_l [64], W_out [16x64], b_out [16]
 * (output rows 3..15 stay 0).  feat: bf16 [n x 16]; out: FP32 [n x 16]. */
typedef struct tlfea_mpnn tlfea_mpnn;
int tlfea_mpnn_create(int layers, tlfea_mpnn** out);
int tlfea_mpnn_destroy(tlfea_mpnn* net);
int tlfea_mpnn_num_params(const tlfea_mpnn* net, int* n_params);
/* host FP32 parameters in, Adam state reset */
int tlfea_mpnn_set_params(tlfea_mpnn* net, const float* host);
int tlfea_mpnn_get_params(const tlfea_mpnn* net, float* host);
int tlfea_mpnn_get_grads(const tlfea_mpnn* net, float* host);
int tlfea_mpnn_zero_grads(tlfea_mpnn* net, void* stream);
int tlfea_mpnn_forward(tlfea_mpnn* net, int n, const unsigned short* feat, const int* off, const int* col,
                       float* out, void* stream);
/* forward + loss sum_i w_i |out_i[:3] - target_i|^2 / sum_i w_i + backward; gradients are
 * added to the gradient buffer; *loss (host) receives the loss. */
int tlfea_mpnn_loss_grad(tlfea_mpnn* net, int n, const unsigned short* feat, const int* off, const int* col,
                         const float* target, const float* weight, double* loss, void* stream);
int tlfea_mpnn_adam(tlfea_mpnn* net, double lr, double beta1, double beta2, double eps, void* stream);
/* Backward of a caller-supplied output gradient dout (FP32 [n x 3]) through the network
 * state of the last tlfea_mpnn_forward on the same n nodes; gradients are added. */
int tlfea_mpnn_backward(tlfea_mpnn* net, int n, const int* off, const int* col, const float* dout, void* stream);
/* Output gradient of the energy loss sum_inst w_inst e^T K e / E_inst with
 * e = scale * out - u*: dout[i, d] = 2 w_over_E[inst] scale[3 inst + d] Ke[3i + d]. */
int tlfea_mpnn_energy_grad(int n, const int* node_inst, const double* scale, const double* Ke,
                           const double* w_over_E, float* dout, void* stream);
/* Features of the nodes of a batch of cantilevers: inst_par[i*12] = bbox low corner (3),
 * extents (3), total load (3), E, nu, unused; fixed_dof_mask and f as in tlfea_batch.
 * Per node: normalised rest coordinates (3), fixed flag, nodal load / |F| (3), F / |F| (3),
 * (log10 E - 8)/2, (nu - 0.325)/0.125, log aspect ratios (2), xi^2, (3 xi^2 - xi^3)/2. */
int tlfea_mpnn_features(int n, const int* node_inst, const double* x, const double* y, const double* z,
                        const unsigned char* fixed_dof_mask, const double* f, const double* inst_par,
                        unsigned short* feat, void* stream);
/* u[3i+d] = scale[3 inst(i) + d] * out[i, d] on free DOFs (0 on fixed ones), the
 * displacement guess (per-instance, per-component scales). */
int tlfea_mpnn_scale_output(int n, const int* node_inst, const float* out, const double* scale,
                            const unsigned char* fixed_dof_mask, double* u, void* stream);

/* ---- mixed precision CG (not in the reference) -------------------------------------- */
/* k32[9*sell_blocks] = (float) k_sell. */
int tlfea_matrix_to_f32(const tlfea_pattern* p, const double* k_sell, float* k32, void* stream);
/* CG on K u = f with the Krylov recursion in FP32 (FP32 block SpMV on k32, FP32 r and p)
 * and FP64 where accuracy is decided: u is accumulated in FP64, dot products and norms are
 * accumulated in FP64, and the true residual f - K u is recomputed with the FP64 matrix
 * (reliable update) whenever the recursive residual has dropped by `refresh` (e.g. 1e-2)
 * since the last refresh or its estimate reaches tol.  Converged when
 * ||f - K u|| / ||f|| < tol in FP64.  rel_res_out is that FP64 ratio; refreshes_out counts
 * the FP64 residual evaluations. */
int tlfea_cg_solve_mixed(tlfea_cg* c, const double* k_sell, const float* k32, const double* f, double* u,
                         double tol, double refresh, int max_iter, int check_every, int* iters_out,
                         double* rel_res_out, int* refreshes_out, void* stream);

/* ---- position update (LinearStaticSolver.cu:189-196,449-463) ----------------------- */
int tlfea_add_displacement(int n_nodes, const double* u, double* x, double* y, double* z, void* stream);


/* ---- implicit-Euler velocity solvers (SyncedNesterov / SyncedAdamW) -------------- */
/* One element body on the device.  The solvers read the element data, update x/y/z in
 * place (as the reference's kernels write d_data->x12() etc.), and use f_int and
 * constraint as workspaces / outputs (their values after a step are those of the last
 * evaluation, like the reference's d_f_int and d_constraint). */
typedef struct tlfea_body {
    int nn, n_elem, n_nodes;
    const int* conn;
    double *x, *y, *z;
    const double *grad_t, *detJ_t;
    tlfea_quad quad;
    tlfea_material mat;
    const tlfea_pattern* pattern; /* node pattern of this mesh */
    const double* m_sell;         /* mass matrix, sliced layout (tlfea_mass_assemble) */
    const double* f_ext;          /* [3*n_nodes] */
    int n_fixed;                  /* constraints = 3 * n_fixed */
    const int* fixed_nodes;       /* [n_fixed] */
    const double *xref, *yref, *zref; /* constraint targets (positions at Setup) */
    double* f_int;                /* [3*n_nodes] */
    double* constraint;           /* [3*n_fixed] */
    const tlfea_force_plan* force_plan; /* optional; NULL = atomic-scatter force kernel */
    /* optional, with force_plan: the positions the gradients were computed from (CalcDnDuPre);
     * when set, f_int is evaluated by tlfea_internal_force_rest's kernel (gradients rebuilt
     * per quadrature point, grad_t not read) */
    const double *X0, *Y0, *Z0;
} tlfea_body;

/* Byte-compatible with the reference's SyncedNesterovParams (SyncedNesterov.cuh:31-35). */
typedef struct tlfea_nesterov_params {
    double alpha, rho, inner_tol, outer_tol;
    int max_outer, max_inner;
    double time_step;
} tlfea_nesterov_params;

/* Byte-compatible with the reference's SyncedAdamWParams (SyncedAdamW.cuh:32-39). */
typedef struct tlfea_adamw_params {
    double lr, beta1, beta2, eps, weight_decay, lr_decay;
    double inner_tol, outer_tol, rho;
    int max_outer, max_inner;
    double time_step;
    int convergence_check_interval;
    double inner_rtol;
} tlfea_adamw_params;

#define TLFEA_SOLVER_NESTEROV 0     /* SyncedNesterovSolver */
#define TLFEA_SOLVER_ADAMW 1        /* SyncedAdamWSolver (cooperative semantics) */
#define TLFEA_SOLVER_ADAMW_NOCOOP 2 /* SyncedAdamWNocoopSolver */

typedef struct tlfea_solver tlfea_solver;
/* Allocates the solver state (v_guess, v_prev, lambda, x_prev, iterates, reduction
 * scratch) for `body`, all zero (SyncedNesterovSolver::Setup, SyncedNesterov.cuh:157-179).
 * The body descriptor is copied; its device arrays must outlive the solver. */
int tlfea_solver_create(int kind, const tlfea_body* body, tlfea_solver** out);
int tlfea_solver_destroy(tlfea_solver* s);
/* Replace the body descriptor (same mesh sizes and constraint count), e.g. after a material
 * change; the solver state is kept. */
int tlfea_solver_update_body(tlfea_solver* s, const tlfea_body* body);
/* SetParameters: stores the parameters and zeroes v_guess, v_prev and lambda
 * (SyncedNesterov.cuh:139-155, SyncedAdamW.cuh:160-190).  params points at a
 * tlfea_nesterov_params or tlfea_adamw_params according to the solver kind. */
int tlfea_solver_set_parameters(tlfea_solver* s, const void* params);
/* Optional relative stop, not in the reference: also end the inner loop once
 * ||g_free|| <= rel_tol * ||g_free at the first inner iteration|| (fixed DOFs excluded).
 * 0 (the default) disables it. */
int tlfea_solver_set_rel_tol(tlfea_solver* s, double rel_tol);
/* Optional momentum restart of the Nesterov solver, not in the reference (whose t-sequence
 * SyncedNesterov.cu:152-175 never restarts, so its iterates converge sublinearly): with
 * on = 1, t restarts at 1 whenever g(y_k).(v_(k+1) - v_k) > 0 (the gradient test of
 * O'Donoghue and Candes, 2015), which makes the inner loop converge linearly on these
 * strongly convex steps.  0 (the default) keeps the reference's sequence.  EINVAL for the
 * AdamW kinds.  tlfea_solver_restarts gives the count of the last Solve. */
int tlfea_solver_set_restart(tlfea_solver* s, int on);
int tlfea_solver_restarts(const tlfea_solver* s, int* restarts);
/* Not in the reference: applies the mass term of the velocity solvers' gradient
 * (M (v_guess - v_prev) / dt, SyncedNesterov.cu:32-95 and the AdamW kinds' equivalent)
 * without the assembled mass: per element, inside the rest-coordinate force kernel, with the
 * quadrature rule of mass_matrix_qp_kernel (FEAT10Data.cu:200-271) and density rho, so
 * the inner iteration reads no mass matrix.  The same operator as a mass assembled with
 * tlfea_mass_assemble at that density (the caller's statement; the body's m_sell is then
 * not read).  rho = 0 returns to the assembled mass.  Needs a body with a force plan and
 * rest positions, SVK without damping, the standard rule (EINVAL otherwise). */
int tlfea_solver_set_matrix_free_mass(tlfea_solver* s, double rho);
/* How often (in inner iterations) the host reads the device convergence flag to stop
 * launching; results do not depend on it.  Default 16. */
int tlfea_solver_set_poll_interval(tlfea_solver* s, int every);
/* One Solve() = one implicit time step (SyncedNesterov.cu:97-360, SyncedAdamW.cu:99-344,
 * SyncedAdamWNocoop.cu:255-458).  Synchronous, like the reference. */
int tlfea_solver_solve(tlfea_solver* s, void* stream);
/* Device pointers of the solver state: v_guess, v_prev [3*n_nodes], lambda [3*n_fixed]. */
int tlfea_solver_state(tlfea_solver* s, double** v_guess, double** v_prev, double** lambda_guess);
/* Statistics of the last Solve(): inner iterations executed (summed over outer
 * iterations), outer iterations executed, last ||g||, last ||g_free||/||g_free,0||. */
int tlfea_solver_stats(const tlfea_solver* s, int* inner_iters, int* outer_iters, double* norm_g, double* rel_g);

#ifdef __cplusplus
}
#endif
#endif /* TLFEA_HIP_H */
