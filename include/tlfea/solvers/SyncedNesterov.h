// tlfea/solvers/SyncedNesterov.h -- implicit-Euler Nesterov velocity solver
// (reference: src/solvers/SyncedNesterov.cuh:31-322).  One Solve() = one time step.
#pragma once

#include "../../tlfea_hip.h"
#include "../elements/ElementBase.h"
#include "SolverBase.h"

namespace tlfea {

// Same fields, order and types as the reference (the SetParameters(void*) payload).
struct SyncedNesterovParams {
    Real alpha, rho, inner_tol, outer_tol;
    int max_outer, max_inner;
    Real time_step;
};

namespace detail {
// Common driver of the three velocity solvers over tlfea_solver (C-ABI).
class VelocitySolverBase : public SolverBase {
  public:
    VelocitySolverBase(ElementBase* data, int n_constraints, int kind);
    ~VelocitySolverBase() override;
    VelocitySolverBase(const VelocitySolverBase&) = delete;
    VelocitySolverBase& operator=(const VelocitySolverBase&) = delete;
    void Setup();
    void Solve() override;
    // not in the reference: statistics of the last Solve() and an optional relative stop
    int GetLastInnerIterations() const { return last_inner_; }
    int GetLastOuterIterations() const { return last_outer_; }
    Real GetLastRelativeGradient() const { return last_rel_; }
    void SetRelativeTolerance(Real rel_tol);
    // device state (v_guess, v_prev [3 n_coef], lambda [n_constraints])
    Real* GetVelocityDevicePtr();
    Real* GetPrevVelocityDevicePtr();
    Real* GetLambdaDevicePtr();

  protected:
    void set_params(const void* p);
    ElementBase* data_;
    int kind_;
    int n_coef_, n_constraints_;
    tlfea_solver* h_ = nullptr;
    bool have_params_ = false;
    unsigned char params_[sizeof(tlfea_adamw_params)];
    Real rel_tol_ = 0.0;
    int last_inner_ = 0, last_outer_ = 0;
    Real last_rel_ = 0.0;
};
}  // namespace detail

class SyncedNesterovSolver : public detail::VelocitySolverBase {
  public:
    SyncedNesterovSolver(ElementBase* data, int n_constraints)
        : VelocitySolverBase(data, n_constraints, TLFEA_SOLVER_NESTEROV) {}
    // Stores the parameters and zeroes v_guess, v_prev and lambda (SyncedNesterov.cuh:139-155).
    void SetParameters(void* params) override { set_params(params); }
    void OneStepNesterov() { Solve(); }
};

}  // namespace tlfea
