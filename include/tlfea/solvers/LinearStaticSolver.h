// tlfea/solvers/LinearStaticSolver.h -- linear static solve K(x) u = f_ext with Dirichlet
// elimination and CG, then x += u (reference: src/solvers/LinearStaticSolver.cuh:38-122).
#pragma once

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
};

using LinearStaticSolverT10 = LinearStaticSolver<GPU_FEAT10_Data>;
using LinearStaticSolverT4 = LinearStaticSolver<GPU_FEAT4_Data>;

extern template class LinearStaticSolver<GPU_FEAT10_Data>;
extern template class LinearStaticSolver<GPU_FEAT4_Data>;

}  // namespace tlfea
