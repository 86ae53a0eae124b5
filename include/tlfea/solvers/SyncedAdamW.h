// tlfea/solvers/SyncedAdamW.h -- implicit-Euler AdamW velocity solver, cooperative
// semantics (reference: src/solvers/SyncedAdamW.cuh:25-387).
#pragma once

#include "SyncedNesterov.h"

namespace tlfea {

struct SyncedAdamWParams {
    Real lr, beta1, beta2, eps, weight_decay, lr_decay;
    Real inner_tol, outer_tol, rho;
    int max_outer, max_inner;
    Real time_step;
    int convergence_check_interval;
    Real inner_rtol;
};

class SyncedAdamWSolver : public detail::VelocitySolverBase {
  public:
    SyncedAdamWSolver(ElementBase* data, int n_constraints)
        : VelocitySolverBase(data, n_constraints, TLFEA_SOLVER_ADAMW) {}
    void SetParameters(void* params) override { set_params(params); }
    void OneStepAdamW() { Solve(); }
};

}  // namespace tlfea
