// tlfea/solvers/SolverBase.h -- solver interface (reference: src/solvers/SolverBase.h:18-25).
#pragma once

namespace tlfea {

class SolverBase {
  public:
    virtual ~SolverBase() = default;
    virtual void Solve() = 0;
    // Parameters travel as a pointer to the solver's POD parameter struct.
    virtual void SetParameters(void* params) = 0;
};

}  // namespace tlfea
