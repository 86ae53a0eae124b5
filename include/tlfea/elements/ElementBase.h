// tlfea/elements/ElementBase.h -- element interface of the drop-in API
// (reference: src/elements/ElementBase.h:22-48).
#pragma once

#include <vector>

#include "../types.h"

namespace tlfea {

enum ElementType { TYPE_3243, TYPE_3443, TYPE_T10, TYPE_T4 };

class ElementBase {
  public:
    ElementType type;

    virtual ~ElementBase() {}

    // The reference keeps a bitwise device copy of the element object here.  This library
    // has no device mirror of the object (kernels take plain device arrays through the
    // C-ABI); the member is kept for source compatibility and points at the object.
    ElementBase* d_data = nullptr;

    virtual int get_n_beam() const = 0;
    virtual int get_n_coef() const = 0;

    virtual void CalcMassMatrix() = 0;
    virtual void CalcInternalForce() = 0;
    virtual void CalcConstraintData() = 0;
    virtual void CalcP() = 0;
    virtual void RetrieveInternalForceToCPU(VectorXR& internal_force) = 0;
    virtual void RetrieveConstraintDataToCPU(VectorXR& constraint) = 0;
    virtual void RetrieveConstraintJacobianToCPU(MatrixXR& constraint_jac) = 0;
    virtual void RetrievePositionToCPU(VectorXR& x12, VectorXR& y12, VectorXR& z12) = 0;
    virtual void RetrieveDeformationGradientToCPU(std::vector<std::vector<MatrixXR>>& deformation_gradient) = 0;
    virtual void RetrievePFromFToCPU(std::vector<std::vector<MatrixXR>>& p_from_F) = 0;
};

}  // namespace tlfea
