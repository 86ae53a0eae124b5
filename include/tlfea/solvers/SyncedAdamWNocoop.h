// tlfea/solvers/SyncedAdamWNocoop.h -- AdamW velocity solver with the host-driven
// semantics of src/solvers/SyncedAdamWNocoop.cuh:25-419 (host bias factors, one outer
// iteration without constraints, outer stop only after a converged inner loop).
#pragma once

#include "SyncedAdamW.h"

namespace tlfea {

using SyncedAdamWNocoopParams = SyncedAdamWParams;

class SyncedAdamWNocoopSolver : public detail::VelocitySolverBase {
  public:
    SyncedAdamWNocoopSolver(ElementBase* data, int n_constraints)
        : VelocitySolverBase(data, n_constraints, TLFEA_SOLVER_ADAMW_NOCOOP) {}
    void SetParameters(void* params) override { set_params(params); }
    void OneStepAdamWNocoop() { Solve(); }
};

}  // namespace tlfea
