// Drop-in solver classes over the C-ABI (tlfea_hip.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../../include/tlfea/solvers/LinearStaticSolver.h"
#include "../../../include/tlfea/solvers/SyncedAdamWNocoop.h"

namespace tlfea {

namespace {
void lib_check(int rc, const char* what) {
    if (rc != TLFEA_OK) throw std::runtime_error(std::string(what) + ": " + tlfea_last_error());
}

detail::TetElementData* as_tet(ElementBase* data) {
    if (data == nullptr) return nullptr;
    if (data->type != TYPE_T10 && data->type != TYPE_T4) return nullptr;
    return static_cast<detail::TetElementData*>(data);
}

// u = x - X0 per coordinate (the body's displacement from its rest positions)
__global__ void k_displacement(int n, const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, const double* __restrict__ X0,
                               const double* __restrict__ Y0, const double* __restrict__ Z0, double* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u[i] = x[i] - X0[i];
    u[n + i] = y[i] - Y0[i];
    u[2 * n + i] = z[i] - Z0[i];
}

static_assert(sizeof(SyncedNesterovParams) == sizeof(tlfea_nesterov_params), "param layout");
static_assert(sizeof(SyncedAdamWParams) == sizeof(tlfea_adamw_params), "param layout");
}  // namespace

namespace detail {

// SyncedNesterov.cuh:39-110: the reference dispatches on the element type and rejects
// unknown types; this library has no ANCF elements.
VelocitySolverBase::VelocitySolverBase(ElementBase* data, int n_constraints, int kind)
    : data_(data), kind_(kind), n_coef_(data ? data->get_n_coef() : 0), n_constraints_(n_constraints) {
    std::memset(params_, 0, sizeof(params_));
    if (!as_tet(data))
        std::fprintf(stderr, "[TLFEA ERROR] unsupported element type for this solver (TET10 and TET4 only)\n");
}

VelocitySolverBase::~VelocitySolverBase() {
    if (h_) tlfea_solver_destroy(h_);
}

// SyncedNesterov.cuh:157-179: (re)allocates zeroed state.  The element's mass matrix must
// have been computed (CalcMassMatrix), as the reference's gradient reads it.
void VelocitySolverBase::Setup() {
    TetElementData* e = as_tet(data_);
    if (!e) return;
    if (h_) {
        tlfea_solver_destroy(h_);
        h_ = nullptr;
    }
    tlfea_body b = e->BodyDescriptor();
    if (3 * b.n_fixed != n_constraints_)
        std::fprintf(stderr,
                     "[TLFEA ERROR] solver built for %d constraints but the element has %d; using the element's\n",
                     n_constraints_, 3 * b.n_fixed);
    lib_check(tlfea_solver_create(kind_, &b, &h_), "solver Setup");
    if (have_params_) lib_check(tlfea_solver_set_parameters(h_, params_), "solver Setup");
    lib_check(tlfea_solver_set_rel_tol(h_, rel_tol_), "solver Setup");
}

void VelocitySolverBase::set_params(const void* p) {
    if (!p) return;
    std::memcpy(params_, p, kind_ == TLFEA_SOLVER_NESTEROV ? sizeof(tlfea_nesterov_params) : sizeof(tlfea_adamw_params));
    have_params_ = true;
    if (!h_) Setup();
    if (h_) lib_check(tlfea_solver_set_parameters(h_, params_), "SetParameters");
}

void VelocitySolverBase::SetRelativeTolerance(Real rel_tol) {
    rel_tol_ = rel_tol;
    if (h_) lib_check(tlfea_solver_set_rel_tol(h_, rel_tol), "SetRelativeTolerance");
}

void VelocitySolverBase::Solve() {
    TetElementData* e = as_tet(data_);
    if (!e) return;
    if (!h_) Setup();
    if (!have_params_) {
        std::fprintf(stderr, "[TLFEA ERROR] SetParameters must be called before Solve\n");
        return;
    }
    tlfea_body b = e->BodyDescriptor();  // picks up material or load changes
    lib_check(tlfea_solver_update_body(h_, &b), "Solve");
    lib_check(tlfea_solver_solve(h_, nullptr), "Solve");
    double ng = 0.0;
    lib_check(tlfea_solver_stats(h_, &last_inner_, &last_outer_, &ng, &last_rel_), "Solve");
}

Real* VelocitySolverBase::GetVelocityDevicePtr() {
    double* p = nullptr;
    if (h_) tlfea_solver_state(h_, &p, nullptr, nullptr);
    return p;
}
Real* VelocitySolverBase::GetPrevVelocityDevicePtr() {
    double* p = nullptr;
    if (h_) tlfea_solver_state(h_, nullptr, &p, nullptr);
    return p;
}
Real* VelocitySolverBase::GetLambdaDevicePtr() {
    double* p = nullptr;
    if (h_) tlfea_solver_state(h_, nullptr, nullptr, &p);
    return p;
}

}  // namespace detail

// ---------------------------------------------------------------------------------------
// LinearStaticSolver (LinearStaticSolver.cu:206-481)
// ---------------------------------------------------------------------------------------
template <typename TData>
LinearStaticSolver<TData>::LinearStaticSolver(TData* data, Real tol, int max_iter)
    : data_(data), n_dof_(3 * data->get_n_coef()), cg_tol_(tol), cg_max_iter_(max_iter) {}

template <typename TData>
LinearStaticSolver<TData>::~LinearStaticSolver() {
    if (batch_) tlfea_batch_destroy(batch_);
    if (disp_) hipFree(disp_);
    if (cg_) tlfea_cg_destroy(cg_);
    if (k_) hipFree(k_);
    if (f_) hipFree(f_);
    if (u_) hipFree(u_);
    if (jinv_) hipFree(jinv_);
    if (mask_) hipFree(mask_);
}

template <typename TData>
void LinearStaticSolver<TData>::Solve() {
    tlfea_body b = data_->BodyDescriptor();
    int64_t nnz = 0, dof = 0, nb = 0;
    lib_check(tlfea_pattern_sizes(b.pattern, &nnz, &dof, &nb), "LinearStaticSolver");
    if (!cg_) {
        lib_check(tlfea_cg_create(b.pattern, &cg_), "LinearStaticSolver");
        if (hipMalloc((void**)&k_, sizeof(double) * 9 * (size_t)nb) != hipSuccess ||
            hipMalloc((void**)&f_, sizeof(double) * (size_t)n_dof_) != hipSuccess ||
            hipMalloc((void**)&u_, sizeof(double) * (size_t)n_dof_) != hipSuccess)
            throw std::runtime_error("LinearStaticSolver: device allocation failed");
    }
    if (hipMemcpy(f_, b.f_ext, sizeof(double) * n_dof_, hipMemcpyDeviceToDevice) != hipSuccess)
        throw std::runtime_error("LinearStaticSolver: copy of f_ext failed");
    int iters = 0;
    double res = 0.0;
    int rc = TLFEA_OK;
    // K(x) applied without assembly when the rest positions are known (with Kelvin-Voigt
    // damping the operator is K + C); a rejected call falls back to the assembled tangent
    const bool damp = b.mat.eta_damp != 0.0 || b.mat.lambda_damp != 0.0;
    last_two_level_ = two_level_ && !assembled_only_ && b.X0 && !damp && b.mat.model == 0 &&
                      SolveTwoLevel(b, &iters, &res, &rc);
    bool matfree = !last_two_level_ && !assembled_only_ && b.force_plan && b.X0;
    if (matfree) {
        if (!jinv_) {
            const size_t nq = (size_t)b.quad.n;
            if (hipMalloc((void**)&jinv_, sizeof(double) * 9 * nq * (size_t)tlfea_padded_elems(b.n_elem)) != hipSuccess ||
                hipMalloc((void**)&mask_, (size_t)n_dof_) != hipSuccess)
                throw std::runtime_error("LinearStaticSolver: device allocation failed");
            lib_check(tlfea_precompute_jinv(b.nn, b.n_elem, b.conn, b.X0, b.Y0, b.Z0, &b.quad, jinv_, nullptr),
                      "LinearStaticSolver");
            lib_check(tlfea_jinv_affine(b.nn, b.n_elem, jinv_, b.detJ_t, &affine_, nullptr), "LinearStaticSolver");
        }
        // the fixed set may have changed since the last Solve (UpdateNodalFixed)
        lib_check(tlfea_fixed_dof_mask(b.n_nodes, b.n_fixed, b.n_fixed > 0 ? b.fixed_nodes : nullptr, mask_, nullptr),
                  "LinearStaticSolver");
        rc = tlfea_cg_solve_matfree(cg_, b.force_plan, b.x, b.y, b.z, b.X0, b.Y0, b.Z0, jinv_, b.detJ_t, affine_,
                                    &b.quad, &b.mat, mask_, f_, u_, cg_tol_, cg_max_iter_, 32, &iters, &res, nullptr);
        if (damp && rc == TLFEA_EINVAL) {
            matfree = false;
            rc = TLFEA_OK;
        }
    }
    last_matfree_ = matfree || last_two_level_;
    if (!matfree && !last_two_level_) {
        // assemble the tangent at the current positions (h = 1), eliminate fixed DOFs
        lib_check(tlfea_stiffness_assemble(b.pattern, b.nn, b.n_elem, b.conn, b.x, b.y, b.z, b.grad_t, b.detJ_t,
                                           &b.quad, &b.mat, 1.0, k_, nullptr),
                  "AssembleLinearStiffness");
        if (b.n_fixed > 0)
            lib_check(tlfea_apply_dirichlet(b.pattern, b.n_fixed, b.fixed_nodes, k_, f_, nullptr),
                      "ApplyDirichletBCs");
        rc = tlfea_cg_solve(cg_, k_, f_, u_, cg_tol_, cg_max_iter_, 32, &iters, &res, nullptr);
    }
    if (rc == TLFEA_ENOTPD) {
        std::fprintf(stderr, "[TLFEA ERROR] LinearStaticSolver CG: non-positive p*Kp detected; aborting.\n");
    } else if (rc != TLFEA_OK && rc != TLFEA_ENOCONV) {
        lib_check(rc, "SolveLinearSystemCG");
    }
    // the reference reports max_iter and the last residual when CG does not converge
    last_iter_count_ = (rc == TLFEA_ENOCONV) ? cg_max_iter_ : iters;
    last_residual_ = res;
    lib_check(tlfea_add_displacement(b.n_nodes, u_, b.x, b.y, b.z, nullptr), "UpdatePositions");
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("LinearStaticSolver: device error");
}

// The two-level path through a one-instance batch: the batch copies the connectivity, rest
// positions, fixed-DOF mask and Lame constants at creation, so it is rebuilt when the fixed
// set (UpdateNodalFixed) or the material changes; the operator follows the current
// positions through the displacement x - X0 (K at rest while x = X0).  Returns false when
// the batch cannot be built (the plain CG then runs).
template <typename TData>
bool LinearStaticSolver<TData>::SolveTwoLevel(const tlfea_body& b, int* iters, double* res, int* rc) {
    const int N = b.n_nodes;
    std::vector<int> fixed(b.n_fixed);
    if (b.n_fixed > 0 &&
        hipMemcpy(fixed.data(), b.fixed_nodes, sizeof(int) * b.n_fixed, hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("LinearStaticSolver: copy of the fixed nodes failed");
    if (!batch_ || fixed != batch_fixed_ || b.mat.lambda != batch_lam_ || b.mat.mu != batch_mu_) {
        if (batch_) {
            tlfea_batch_destroy(batch_);
            batch_ = nullptr;
        }
        if (!mask_ && hipMalloc((void**)&mask_, (size_t)n_dof_) != hipSuccess)
            throw std::runtime_error("LinearStaticSolver: device allocation failed");
        lib_check(tlfea_fixed_dof_mask(N, b.n_fixed, b.n_fixed > 0 ? b.fixed_nodes : nullptr, mask_, nullptr),
                  "LinearStaticSolver");
        const int eo[2] = {0, b.n_elem}, no[2] = {0, N};
        double lam = b.mat.lambda, mu = b.mat.mu;
        if (tlfea_batch_create(b.nn, 1, eo, no, b.conn, b.X0, b.Y0, b.Z0, &lam, &mu, mask_, &batch_, nullptr) !=
            TLFEA_OK) {
            batch_ = nullptr;
            return false;
        }
        batch_fixed_ = fixed;
        batch_lam_ = b.mat.lambda;
        batch_mu_ = b.mat.mu;
    }
    if (!disp_ && hipMalloc((void**)&disp_, sizeof(double) * 3 * (size_t)N) != hipSuccess)
        throw std::runtime_error("LinearStaticSolver: device allocation failed");
    k_displacement<<<(N + 255) / 256, 256>>>(N, b.x, b.y, b.z, b.X0, b.Y0, b.Z0, disp_);
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("LinearStaticSolver: displacement kernel failed");
    lib_check(tlfea_batch_set_displacement(batch_, disp_, disp_ + N, disp_ + 2 * (size_t)N, nullptr),
              "LinearStaticSolver");
    lib_check(tlfea_batch_precond_setup(batch_, tl_degree_, tl_ratio_, 20, 1.1, nullptr, nullptr),
              "LinearStaticSolver");
    lib_check(tlfea_batch_coarse_setup(batch_, 1, nullptr, nullptr), "LinearStaticSolver");
    int it = 0, st = 0, total = 0;
    double rr = 0.0;
    const int r = tlfea_batch_cg_solve(batch_, f_, u_, 0, cg_tol_, cg_max_iter_, 16, 0, 1, &it, &rr, &st, &total,
                                       nullptr);
    if (r != TLFEA_OK && r != TLFEA_ENOCONV) lib_check(r, "SolveLinearSystemCG");
    *iters = it;
    *res = rr;
    *rc = st;
    return true;
}

template class LinearStaticSolver<GPU_FEAT10_Data>;
template class LinearStaticSolver<GPU_FEAT4_Data>;

}  // namespace tlfea
