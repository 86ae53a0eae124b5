// tlfea/solvers/LinearStaticSolver.h -- linear static solve K(x) u = f_ext with Dirichlet
// elimination and CG, then x += u (reference: src/solvers/LinearStaticSolver.cuh:38-122).
#pragma once

#include <vector>

#include "../../tlfea_hip.h"
#include "../elements/FEAT10Data.h"
#include "../elements/FEAT4Data.h"
#include "SolverBase.h"

namespace tlfea {

template <typename TData>
class LinearStaticSolver : public SolverBase {
  public:
    // The caller has run Setup(), SetSVK(), CalcDnDuPre() and optionally SetNodalFixed() /
    // SetExternalForce().  tol is the CG relative residual sqrt(r.r / r0.r0).
    LinearStaticSolver(TData* data, Real tol = 1e-10, int max_iter = 10000);
    ~LinearStaticSolver() override;
    LinearStaticSolver(const LinearStaticSolver&) = delete;
    LinearStaticSolver& operator=(const LinearStaticSolver&) = delete;

    void Solve() override;
    void SetParameters(void*) override {}
    int GetLastIterCount() const { return last_iter_count_; }
    Real GetLastResidual() const { return last_residual_; }
    // Not in the reference.  Solve applies K(x) without assembling it (the same operator
    // and CG stopping rule, ~5x faster on large meshes) whenever the element's rest
    // positions are known (CalcDnDuPre), with Kelvin-Voigt damping as K + C; otherwise, or with
    // UseAssembledOperator(true), it assembles K (+ C) as the reference does.
    void UseAssembledOperator(bool on) { assembled_only_ = on; }
    bool LastSolveWasMatrixFree() const { return last_matfree_; }
    // Not in the reference.  With on, Solve runs the flexible FP64 CG preconditioned by the
    // FP32 two-level stage (tlfea_batch_precond_setup with `degree` Chebyshev steps on
    // [lambda_max / ratio, lambda_max], and tlfea_batch_coarse_setup) on K(x), set up afresh
    // for every Solve, with the same stopping rule sqrt(r.r / r0.r0) < tol: the same solution
    // in a small fraction of the iterations on slender bodies (80 vs 1,226 on the 20x2x2-cell
    // beam, 102 vs 3,740 on the reference's beam_highres_t10 mesh; 0.26 s then 0.09 s per
    // Solve on a 2M-DOF beam against 0.63 and 0.85 s).  The first Solve also builds the
    // one-instance batch (its force plan and coarse geometry).  SVK bodies
    // without Kelvin-Voigt damping whose rest positions are known (CalcDnDuPre); others keep
    // the plain CG.
    void UseTwoLevelPreconditioner(bool on, int degree = 16, Real ratio = 1000.0) {
        two_level_ = on;
        tl_degree_ = degree;
        tl_ratio_ = ratio;
    }
    bool LastSolveWasTwoLevel() const { return last_two_level_; }

  private:
    TData* data_;
    int n_dof_;
    Real cg_tol_;
    int cg_max_iter_;
    int last_iter_count_ = 0;
    Real last_residual_ = 0.0;
    tlfea_cg* cg_ = nullptr;
    double *k_ = nullptr, *f_ = nullptr, *u_ = nullptr;
    // matrix-free operator: inverse Jacobians at the rest positions, fixed-DOF mask
    double* jinv_ = nullptr;
    unsigned char* mask_ = nullptr;
    int affine_ = 0;
    bool assembled_only_ = false, last_matfree_ = false;
    // two-level path: a one-instance batch of the body (rebuilt when its fixed set or
    // material changes) and the displacement x - X0 that moves its operator to K(x)
    bool two_level_ = false, last_two_level_ = false;
    int tl_degree_ = 16;
    Real tl_ratio_ = 1000.0;
    tlfea_batch* batch_ = nullptr;
    std::vector<int> batch_fixed_;
    double batch_lam_ = 0.0, batch_mu_ = 0.0;
    double* disp_ = nullptr;  // [3][n_nodes] SoA
    bool SolveTwoLevel(const tlfea_body& b, int* iters, double* res, int* rc);
};

using LinearStaticSolverT10 = LinearStaticSolver<GPU_FEAT10_Data>;
using LinearStaticSolverT4 = LinearStaticSolver<GPU_FEAT4_Data>;

extern template class LinearStaticSolver<GPU_FEAT10_Data>;
extern template class LinearStaticSolver<GPU_FEAT4_Data>;

}  // namespace tlfea
