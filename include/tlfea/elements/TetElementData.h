// tlfea/elements/TetElementData.h -- shared implementation of the drop-in TET10 / TET4
// element classes (GPU_FEAT10_Data, GPU_FEAT4_Data).  Public methods and their call-order
// rules follow src/elements/FEAT10Data.cuh:305-731 and src/elements/FEAT4Data.cuh; all
// device work goes through the C-ABI of tlfea_hip.h.
//
// Error behaviour: misuse (wrong call order, size mismatch) prints "[TLFEA ERROR] ..." to
// stderr and returns, as the reference's MOPHI_ERROR-then-return sites do; HIP failures and
// library errors throw std::runtime_error.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../tlfea_hip.h"
#include "../types.h"
#include "ElementBase.h"

namespace tlfea {
namespace detail {

struct TetState;  // device buffers and flags (defined in the library)

class TetElementData : public ElementBase {
  public:
    TetElementData(int nn, int num_elements, int num_nodes);
    ~TetElementData() override;
    TetElementData(const TetElementData&) = delete;
    TetElementData& operator=(const TetElementData&) = delete;

    int get_n_elem() const { return n_elem; }
    int get_n_coef() const override { return n_coef; }
    int get_n_constraint() const { return n_constraint; }
    int get_n_beam() const override { return n_elem; }

    void Initialize();
    void Setup(const VectorXR& qx, const VectorXR& qy, const VectorXR& qz, const VectorXR& qw, const VectorXR& h_x12,
               const VectorXR& h_y12, const VectorXR& h_z12, const MatrixXi& element_connectivity);
    void Destroy();

    void SetDensity(Real rho0);
    void SetDamping(Real eta_damp, Real lambda_damp);
    void SetSVK();
    void SetSVK(Real E, Real nu);
    void SetMooneyRivlin(Real mu10, Real mu01, Real kappa);
    void SetExternalForce(const VectorXR& h_f_ext);
    void SetNodalFixed(const VectorXi& fixed_nodes);
    void UpdateNodalFixed(const VectorXi& fixed_nodes);
    void UpdatePositions(const VectorXR& h_x12, const VectorXR& h_y12, const VectorXR& h_z12);
    void UpdateConstraintTargets(const VectorXR& h_x12, const VectorXR& h_y12, const VectorXR& h_z12);

    void CalcDnDuPre();
    void CalcMassMatrix() override;
    void BuildMassCSRPattern();
    void CalcP() override;
    void CalcInternalForce() override;
    void CalcConstraintData() override;
    void ConvertToCSR_ConstraintJacT();
    void BuildConstraintJacobianTransposeCSR() { ConvertToCSR_ConstraintJacT(); }
    void ConvertToCSR_ConstraintJac();
    void BuildConstraintJacobianCSR() { ConvertToCSR_ConstraintJac(); }

    void RetrieveMassCSRToCPU(std::vector<int>& offsets, std::vector<int>& columns, std::vector<Real>& values);
    void RetrieveInternalForceToCPU(VectorXR& internal_force) override;
    void RetrieveExternalForceToCPU(VectorXR& external_force);
    // The reference leaves these three as no-ops (FEAT10Data.cuh:618-627); so do these.
    void RetrieveConstraintDataToCPU(VectorXR&) override {}
    void RetrieveConstraintJacobianToCPU(MatrixXR&) override {}
    void RetrieveDeformationGradientToCPU(std::vector<std::vector<MatrixXR>>&) override {}
    void RetrievePositionToCPU(VectorXR& x12, VectorXR& y12, VectorXR& z12) override;
    void RetrievePFromFToCPU(std::vector<std::vector<MatrixXR>>& p_from_F) override;
    void RetrieveDnDuPreToCPU(std::vector<std::vector<MatrixXR>>& dn_du_pre);
    void RetrieveDetJToCPU(std::vector<std::vector<Real>>& detJ);
    void RetrieveConnectivityToCPU(MatrixXi& connectivity);
    void WriteOutputVTK(const std::string& filename);

    const Real* GetX12DevicePtr() const;
    const Real* GetY12DevicePtr() const;
    const Real* GetZ12DevicePtr() const;
    Real* GetExternalForceDevicePtr();
    const Real* GetExternalForceDevicePtr() const;
    Real* Get_Constraint_Ptr();
    bool Get_Is_Constraint_Setup();

    // ---- not in the reference: access for the library's solver classes ----
    // Device descriptor of the element (positions, gradients, mass, loads, constraints).
    // Builds the mass pattern / force plan on first use; CalcMassMatrix must have run.
    tlfea_body BodyDescriptor();
    const tlfea_pattern* Pattern();
    // Host copy of the constraint Jacobian CSR that ConvertToCSR_ConstraintJac (J, rows =
    // constraints) or ConvertToCSR_ConstraintJacT (J^T, rows = DOFs) built; the reference
    // exposes these arrays to its kernels only (FEAT10Data.cuh:262-283).  Empty when the
    // CSR has not been built since the last SetNodalFixed / UpdateNodalFixed.
    void RetrieveConstraintJacobianCSRToCPU(bool transpose, std::vector<int>& offsets, std::vector<int>& columns,
                                            std::vector<Real>& values);
    int nodes_per_element() const { return nn_; }
    void* Stream() const { return nullptr; }

    int n_elem;
    int n_coef;
    int n_constraint = 0;

  protected:
    int nn_, nq_;
    std::unique_ptr<TetState> s_;
    bool is_setup = false;
    bool is_constraints_setup = false;
    bool is_csr_setup = false;
    bool is_cj_csr_setup = false;
    bool is_j_csr_setup = false;
};

}  // namespace detail
}  // namespace tlfea
