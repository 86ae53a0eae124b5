// tlfea/FEASolver.h -- name -> object registry of a simulation (reference: src/FEASolver.h:35-138).
// Non-owning: registered elements and solvers must outlive the registry.  Duplicate names
// and null pointers throw std::invalid_argument; unknown names throw std::out_of_range.
#pragma once

#include <map>
#include <stdexcept>
#include <string>

#include "elements/ElementBase.h"
#include "solvers/SolverBase.h"

namespace tlfea {

class FEASolver {
  public:
    FEASolver() = default;
    ~FEASolver() = default;
    FEASolver(const FEASolver&) = delete;
    FEASolver& operator=(const FEASolver&) = delete;
    FEASolver(FEASolver&&) = delete;
    FEASolver& operator=(FEASolver&&) = delete;

    void AddElement(const std::string& name, ElementBase* element) { add(elements_, "element", name, element); }
    ElementBase* GetElement(const std::string& name) { return find(elements_, "element", name); }
    const ElementBase* GetElement(const std::string& name) const { return find(elements_, "element", name); }
    int GetNumElements() const { return static_cast<int>(elements_.size()); }

    void AddSolver(const std::string& name, SolverBase* solver) { add(solvers_, "solver", name, solver); }
    SolverBase* GetSolver(const std::string& name) { return find(solvers_, "solver", name); }
    const SolverBase* GetSolver(const std::string& name) const { return find(solvers_, "solver", name); }
    int GetNumSolvers() const { return static_cast<int>(solvers_.size()); }

  private:
    template <typename T>
    static void add(std::map<std::string, T*>& m, const char* what, const std::string& name, T* p) {
        const std::string fn = std::string("FEASolver::Add") + (what[0] == 'e' ? "Element" : "Solver");
        if (p == nullptr) throw std::invalid_argument(fn + ": " + what + " pointer is null");
        if (!m.emplace(name, p).second)
            throw std::invalid_argument(fn + ": " + what + " '" + name + "' already registered");
    }
    template <typename T>
    static T* find(const std::map<std::string, T*>& m, const char* what, const std::string& name) {
        auto it = m.find(name);
        if (it == m.end()) {
            const std::string fn = std::string("FEASolver::Get") + (what[0] == 'e' ? "Element" : "Solver");
            throw std::out_of_range(fn + ": " + what + " '" + name + "' not found");
        }
        return it->second;
    }
    std::map<std::string, ElementBase*> elements_;
    std::map<std::string, SolverBase*> solvers_;
};

}  // namespace tlfea
