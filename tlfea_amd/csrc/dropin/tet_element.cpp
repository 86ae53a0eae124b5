// Drop-in TET10 / TET4 element classes: host C++ over the C-ABI (tlfea_hip.h).
// Method semantics follow src/elements/FEAT10Data.cuh / FEAT10Data.cu (cited per method).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/tlfea/elements/TetElementData.h"

namespace tlfea {
namespace detail {

namespace {
void report(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "[TLFEA ERROR] ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
    va_end(ap);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

void lib_check(int rc, const char* what) {
    if (rc != TLFEA_OK) throw std::runtime_error(std::string(what) + ": " + tlfea_last_error());
}

template <typename T>
T* dalloc(size_t n) {
    T* p = nullptr;
    hip_check(hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)), "hipMalloc");
    hip_check(hipMemset(p, 0, sizeof(T) * (n ? n : 1)), "hipMemset");
    return p;
}

template <typename T>
void dfree(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

template <typename T>
void h2d(T* d, const T* h, size_t n) {
    if (n) hip_check(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

template <typename T>
void d2h(T* h, const T* d, size_t n) {
    if (n) hip_check(hipMemcpy(h, d, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
}  // namespace

struct TetState {
    // device arrays
    int* conn = nullptr;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *xj = nullptr, *yj = nullptr, *zj = nullptr;  // constraint targets (*_jac)
    double *x0 = nullptr, *y0 = nullptr, *z0 = nullptr;  // positions at the last CalcDnDuPre
    bool have_rest = false;
    double *grad_t = nullptr, *detJ_t = nullptr;
    double *F_t = nullptr, *P_t = nullptr, *Fd_t = nullptr, *Pv_t = nullptr;
    double *f_int = nullptr, *f_ext = nullptr;
    double* constraint = nullptr;
    int* fixed = nullptr;
    double* m_sell = nullptr;
    // constraint Jacobians (J and J^T, CSR)
    int *j_off = nullptr, *j_col = nullptr, *jt_off = nullptr, *jt_col = nullptr;
    double *j_val = nullptr, *jt_val = nullptr;
    tlfea_pattern* pattern = nullptr;
    tlfea_force_plan* plan = nullptr;
    bool plan_tried = false;
    tlfea_quad quad{};
    double rho0 = 0.0, E = 0.0, nu = 0.0, eta = 0.0, lambda_d = 0.0;
    tlfea_material mat{0.0, 0.0, 0.0, 0.0};
    int material_model = 0;
    bool initialized = false;
    int n_fixed_alloc = 0;
};

TetElementData::TetElementData(int nn, int num_elements, int num_nodes)
    : n_elem(num_elements), n_coef(num_nodes), nn_(nn), nq_(nn == 10 ? 5 : 1), s_(new TetState()) {}

TetElementData::~TetElementData() {
    // The reference frees only in Destroy() (FEAT10Data.cuh:659-723); releasing here as well
    // keeps a forgotten Destroy() from leaking device memory.
    if (s_ && s_->initialized) Destroy();
}

// FEAT10Data.cuh:358-416: per-node and per-element arrays.
void TetElementData::Initialize() {
    TetState& s = *s_;
    if (s.initialized) return;
    const size_t ne_pad = (size_t)tlfea_padded_elems(n_elem);
    s.conn = dalloc<int>((size_t)n_elem * nn_);
    s.x = dalloc<double>(n_coef);
    s.y = dalloc<double>(n_coef);
    s.z = dalloc<double>(n_coef);
    s.xj = dalloc<double>(n_coef);
    s.yj = dalloc<double>(n_coef);
    s.zj = dalloc<double>(n_coef);
    s.x0 = dalloc<double>(n_coef);
    s.y0 = dalloc<double>(n_coef);
    s.z0 = dalloc<double>(n_coef);
    s.grad_t = dalloc<double>(ne_pad * nq_ * nn_ * 3);
    s.detJ_t = dalloc<double>(ne_pad * nq_);
    s.F_t = dalloc<double>(ne_pad * nq_ * 9);
    s.P_t = dalloc<double>(ne_pad * nq_ * 9);
    s.Fd_t = dalloc<double>(ne_pad * nq_ * 9);
    s.Pv_t = dalloc<double>(ne_pad * nq_ * 9);
    s.f_int = dalloc<double>(3 * (size_t)n_coef);
    s.f_ext = dalloc<double>(3 * (size_t)n_coef);
    s.initialized = true;
}

// FEAT10Data.cuh:418-503: upload positions (also as constraint targets), connectivity and
// quadrature; zero gradients, stresses and f_int; material scalars start at zero.
void TetElementData::Setup(const VectorXR& qx, const VectorXR& qy, const VectorXR& qz, const VectorXR& qw,
                           const VectorXR& h_x12, const VectorXR& h_y12, const VectorXR& h_z12,
                           const MatrixXi& element_connectivity) {
    if (is_setup) {
        report("%s is already set up.", nn_ == 10 ? "GPU_FEAT10_Data" : "GPU_FEAT4_Data");
        return;
    }
    if (!s_->initialized) Initialize();
    if (qx.size() != nq_ || qy.size() != nq_ || qz.size() != nq_ || qw.size() != nq_) {
        report("quadrature rule must have %d points", nq_);
        return;
    }
    if (h_x12.size() != n_coef || h_y12.size() != n_coef || h_z12.size() != n_coef ||
        element_connectivity.rows() != n_elem || element_connectivity.cols() != nn_) {
        report("Setup: position or connectivity size mismatch");
        return;
    }
    for (std::ptrdiff_t k = 0; k < element_connectivity.size(); k++) {
        const int v = element_connectivity.data()[k];
        if (v < 0 || v >= n_coef) {
            report("Setup: connectivity entry %d outside [0, %d)", v, n_coef);
            return;
        }
    }
    TetState& s = *s_;
    h2d(s.x, h_x12.data(), n_coef);
    h2d(s.y, h_y12.data(), n_coef);
    h2d(s.z, h_z12.data(), n_coef);
    h2d(s.xj, h_x12.data(), n_coef);
    h2d(s.yj, h_y12.data(), n_coef);
    h2d(s.zj, h_z12.data(), n_coef);
    h2d(s.conn, element_connectivity.data(), (size_t)n_elem * nn_);  // column-major like the reference
    s.quad.n = nq_;
    for (int q = 0; q < nq_; q++) {
        s.quad.x[q] = qx(q);
        s.quad.y[q] = qy(q);
        s.quad.z[q] = qz(q);
        s.quad.w[q] = qw(q);
    }
    s.rho0 = s.E = s.nu = s.eta = s.lambda_d = 0.0;
    s.mat = tlfea_material{0.0, 0.0, 0.0, 0.0};
    s.material_model = 0;
    is_setup = true;
}

void TetElementData::SetDensity(Real rho0) {
    if (!is_setup) {
        report("element must be set up before setting density.");
        return;
    }
    s_->rho0 = rho0;
}

void TetElementData::SetDamping(Real eta_damp, Real lambda_damp) {
    if (!is_setup) {
        report("element must be set up before setting damping.");
        return;
    }
    s_->eta = eta_damp;
    s_->lambda_d = lambda_damp;
    s_->mat.eta_damp = eta_damp;
    s_->mat.lambda_damp = lambda_damp;
}

// SVK with the stored lambda, mu; the Mooney-Rivlin coefficients are cleared
// (FEAT10Data.cuh:533-552)
void TetElementData::SetSVK() {
    if (!is_setup) {
        report("element must be set up before setting material.");
        return;
    }
    s_->material_model = 0;
    s_->mat.model = 0;
    s_->mat.mu10 = s_->mat.mu01 = s_->mat.kappa = 0.0;
}

// mu = E / (2(1+nu)), lambda = E nu / ((1+nu)(1-2nu))  (FEAT10Data.cuh:532-553)
void TetElementData::SetSVK(Real E, Real nu) {
    if (!is_setup) {
        report("element must be set up before setting material.");
        return;
    }
    s_->E = E;
    s_->nu = nu;
    s_->mat.mu = E / (2 * (1 + nu));
    s_->mat.lambda = (E * nu) / ((1 + nu) * (1 - 2 * nu));
    SetSVK();
}

// Compressible Mooney-Rivlin: every later P, internal force and tangent uses
// W = mu10 (J^-2/3 I1 - 3) + mu01 (J^-4/3 I2 - 3) + kappa/2 (J - 1)^2 (FEAT10Data.cuh:571-587)
void TetElementData::SetMooneyRivlin(Real mu10, Real mu01, Real kappa) {
    if (!is_setup) {
        report("element must be set up before setting material.");
        return;
    }
    s_->material_model = 1;
    s_->mat.model = 1;
    s_->mat.mu10 = mu10;
    s_->mat.mu01 = mu01;
    s_->mat.kappa = kappa;
}

void TetElementData::SetExternalForce(const VectorXR& h_f_ext) {
    if (h_f_ext.size() != 3 * (std::ptrdiff_t)n_coef) {
        report("External force vector size mismatch.");
        return;
    }
    if (!s_->initialized) Initialize();
    h2d(s_->f_ext, h_f_ext.data(), 3 * (size_t)n_coef);
}

// FEAT10Data.cu:650-672
void TetElementData::SetNodalFixed(const VectorXi& fixed_nodes) {
    if (is_constraints_setup) {
        report("element CONSTRAINT is already set up.");
        return;
    }
    for (std::ptrdiff_t k = 0; k < fixed_nodes.size(); k++)
        if (fixed_nodes(k) < 0 || fixed_nodes(k) >= n_coef) {
            report("fixed node %d outside [0, %d)", fixed_nodes(k), n_coef);
            return;
        }
    TetState& s = *s_;
    n_constraint = (int)fixed_nodes.size() * 3;
    dfree(s.constraint);
    dfree(s.fixed);
    s.constraint = dalloc<double>(n_constraint);
    s.fixed = dalloc<int>(fixed_nodes.size());
    h2d(s.fixed, fixed_nodes.data(), fixed_nodes.size());
    s.n_fixed_alloc = (int)fixed_nodes.size();
    is_constraints_setup = true;
}

// FEAT10Data.cu:674-751: replace the fixed list, clear c, drop the J / J^T CSR caches.
void TetElementData::UpdateNodalFixed(const VectorXi& fixed_nodes) {
    if (!is_constraints_setup) {
        SetNodalFixed(fixed_nodes);
        return;
    }
    is_constraints_setup = false;
    TetState& s = *s_;
    dfree(s.j_off);
    dfree(s.j_col);
    dfree(s.j_val);
    dfree(s.jt_off);
    dfree(s.jt_col);
    dfree(s.jt_val);
    is_cj_csr_setup = is_j_csr_setup = false;
    SetNodalFixed(fixed_nodes);
}

void TetElementData::UpdatePositions(const VectorXR& h_x12, const VectorXR& h_y12, const VectorXR& h_z12) {
    if (h_x12.size() != n_coef || h_y12.size() != n_coef || h_z12.size() != n_coef) {
        report("Position vector size mismatch.");
        return;
    }
    h2d(s_->x, h_x12.data(), n_coef);
    h2d(s_->y, h_y12.data(), n_coef);
    h2d(s_->z, h_z12.data(), n_coef);
}

void TetElementData::UpdateConstraintTargets(const VectorXR& h_x12, const VectorXR& h_y12, const VectorXR& h_z12) {
    if (h_x12.size() != n_coef || h_y12.size() != n_coef || h_z12.size() != n_coef) {
        report("Position vector size mismatch.");
        return;
    }
    h2d(s_->xj, h_x12.data(), n_coef);
    h2d(s_->yj, h_y12.data(), n_coef);
    h2d(s_->zj, h_z12.data(), n_coef);
}

// FEAT10Data.cu:277-285
void TetElementData::CalcDnDuPre() {
    if (!is_setup) {
        report("CalcDnDuPre before Setup");
        return;
    }
    TetState& s = *s_;
    lib_check(tlfea_precompute_grad(nn_, n_elem, s.conn, s.x, s.y, s.z, &s.quad, s.grad_t, s.detJ_t, nullptr),
              "CalcDnDuPre");
    // the positions the gradients come from, for the force kernel that rebuilds them
    for (int d = 0; d < 3; d++)
        hip_check(hipMemcpy(d == 0 ? s.x0 : d == 1 ? s.y0 : s.z0, d == 0 ? s.x : d == 1 ? s.y : s.z,
                            sizeof(double) * n_coef, hipMemcpyDeviceToDevice),
                  "CalcDnDuPre");
    s.have_rest = true;
    hip_check(hipDeviceSynchronize(), "CalcDnDuPre");
}

// FEAT10Data.cu:361-417 (pattern built once)
void TetElementData::BuildMassCSRPattern() {
    if (is_csr_setup) return;
    TetState& s = *s_;
    lib_check(tlfea_pattern_create(nn_, n_elem, n_coef, s.conn, &s.pattern, nullptr), "BuildMassCSRPattern");
    int64_t nnz = 0, dof = 0, nb = 0;
    lib_check(tlfea_pattern_sizes(s.pattern, &nnz, &dof, &nb), "BuildMassCSRPattern");
    s.m_sell = dalloc<double>(nb);
    is_csr_setup = true;
}

// FEAT10Data.cu:341-359
void TetElementData::CalcMassMatrix() {
    if (!is_csr_setup) BuildMassCSRPattern();
    TetState& s = *s_;
    lib_check(tlfea_mass_assemble(s.pattern, nn_, n_elem, s.conn, s.detJ_t, &s.quad, s.rho0, s.m_sell, nullptr),
              "CalcMassMatrix");
    hip_check(hipDeviceSynchronize(), "CalcMassMatrix");
}

// FEAT10Data.cu:287-305: F, P (no damping: no solver velocity), Fdot = P_vis = 0.
void TetElementData::CalcP() {
    TetState& s = *s_;
    lib_check(tlfea_calc_p(nn_, n_elem, s.conn, s.x, s.y, s.z, s.grad_t, &s.mat, nullptr, s.F_t, s.P_t, s.Fd_t,
                           s.Pv_t, nullptr),
              "CalcP");
    hip_check(hipDeviceSynchronize(), "CalcP");
}

// FEAT10Data.cu:307-323: accumulates onto f_int (only Setup zeroes it).
void TetElementData::CalcInternalForce() {
    TetState& s = *s_;
    lib_check(tlfea_fint_from_p(nn_, n_elem, s.conn, s.P_t, s.grad_t, s.detJ_t, &s.quad, s.f_int, nullptr),
              "CalcInternalForce");
    hip_check(hipDeviceSynchronize(), "CalcInternalForce");
}

// FEAT10Data.cu:325-339
void TetElementData::CalcConstraintData() {
    if (!is_constraints_setup) {
        report("constraint is not set up");
        return;
    }
    if (n_constraint == 0) return;
    TetState& s = *s_;
    lib_check(tlfea_constraint_eval(n_constraint / 3, s.fixed, s.x, s.y, s.z, s.xj, s.yj, s.zj, s.constraint, nullptr),
              "CalcConstraintData");
    hip_check(hipDeviceSynchronize(), "CalcConstraintData");
}

// J: rows = constraints, one entry 1.0 at DOF 3*fixed[k]+d (FEAT10Data.cu:420-506)
void TetElementData::ConvertToCSR_ConstraintJac() {
    if (is_j_csr_setup) return;
    if (!is_constraints_setup || n_constraint == 0) return;
    TetState& s = *s_;
    std::vector<int> fixed(n_constraint / 3);
    d2h(fixed.data(), s.fixed, fixed.size());
    std::vector<int> off(n_constraint + 1), col(n_constraint);
    std::vector<double> val(n_constraint, 1.0);
    for (int t = 0; t < n_constraint; t++) {
        off[t] = t;
        col[t] = 3 * fixed[t / 3] + t % 3;
    }
    off[n_constraint] = n_constraint;
    s.j_off = dalloc<int>(off.size());
    s.j_col = dalloc<int>(col.size());
    s.j_val = dalloc<double>(val.size());
    h2d(s.j_off, off.data(), off.size());
    h2d(s.j_col, col.data(), col.size());
    h2d(s.j_val, val.data(), val.size());
    is_j_csr_setup = true;
}

// J^T: rows = 3N DOFs, columns = constraint index (FEAT10Data.cu:510-562).  Columns of a
// row are ascending here (the reference's atomic fill leaves them in arrival order).
void TetElementData::ConvertToCSR_ConstraintJacT() {
    if (is_cj_csr_setup) return;
    if (!is_constraints_setup || n_constraint == 0) return;
    TetState& s = *s_;
    std::vector<int> fixed(n_constraint / 3);
    d2h(fixed.data(), s.fixed, fixed.size());
    const int rows = 3 * n_coef;
    std::vector<int> off(rows + 1, 0), col(n_constraint);
    for (int t = 0; t < n_constraint; t++) off[3 * fixed[t / 3] + t % 3 + 1]++;
    for (int r = 0; r < rows; r++) off[r + 1] += off[r];
    std::vector<int> cur(off.begin(), off.end() - 1);
    for (int t = 0; t < n_constraint; t++) col[cur[3 * fixed[t / 3] + t % 3]++] = t;
    std::vector<double> val(n_constraint, 1.0);
    s.jt_off = dalloc<int>(off.size());
    s.jt_col = dalloc<int>(col.size());
    s.jt_val = dalloc<double>(val.size());
    h2d(s.jt_off, off.data(), off.size());
    h2d(s.jt_col, col.data(), col.size());
    h2d(s.jt_val, val.data(), val.size());
    is_cj_csr_setup = true;
}

void TetElementData::RetrieveConstraintJacobianCSRToCPU(bool transpose, std::vector<int>& offsets,
                                                       std::vector<int>& columns, std::vector<Real>& values) {
    offsets.clear();
    columns.clear();
    values.clear();
    if (!(transpose ? is_cj_csr_setup : is_j_csr_setup)) return;
    TetState& s = *s_;
    const size_t rows = transpose ? 3 * (size_t)n_coef : (size_t)n_constraint;
    offsets.resize(rows + 1);
    d2h(offsets.data(), transpose ? s.jt_off : s.j_off, rows + 1);
    const size_t nnz = (size_t)offsets[rows];
    columns.resize(nnz);
    values.resize(nnz);
    d2h(columns.data(), transpose ? s.jt_col : s.j_col, nnz);
    d2h(values.data(), transpose ? s.jt_val : s.j_val, nnz);
}

void TetElementData::RetrieveMassCSRToCPU(std::vector<int>& offsets, std::vector<int>& columns,
                                          std::vector<Real>& values) {
    offsets.assign((size_t)n_coef + 1, 0);
    columns.clear();
    values.clear();
    if (!is_csr_setup) return;
    TetState& s = *s_;
    int64_t nnz = 0;
    lib_check(tlfea_pattern_sizes(s.pattern, &nnz, nullptr, nullptr), "RetrieveMassCSRToCPU");
    int *d_off = dalloc<int>(n_coef + 1), *d_col = dalloc<int>(nnz);
    double* d_val = dalloc<double>(nnz);
    lib_check(tlfea_pattern_node_csr(s.pattern, d_off, d_col, nullptr), "RetrieveMassCSRToCPU");
    lib_check(tlfea_mass_to_csr(s.pattern, s.m_sell, d_val, nullptr), "RetrieveMassCSRToCPU");
    columns.resize(nnz);
    values.resize(nnz);
    d2h(offsets.data(), d_off, n_coef + 1);
    d2h(columns.data(), d_col, nnz);
    d2h(values.data(), d_val, nnz);
    dfree(d_off);
    dfree(d_col);
    dfree(d_val);
}

void TetElementData::RetrieveInternalForceToCPU(VectorXR& internal_force) {
    internal_force.resize(3 * n_coef);
    d2h(internal_force.data(), s_->f_int, 3 * (size_t)n_coef);
}

void TetElementData::RetrieveExternalForceToCPU(VectorXR& external_force) {
    external_force.resize(3 * n_coef);
    d2h(external_force.data(), s_->f_ext, 3 * (size_t)n_coef);
}

void TetElementData::RetrievePositionToCPU(VectorXR& x12, VectorXR& y12, VectorXR& z12) {
    x12.resize(n_coef);
    y12.resize(n_coef);
    z12.resize(n_coef);
    d2h(x12.data(), s_->x, n_coef);
    d2h(y12.data(), s_->y, n_coef);
    d2h(z12.data(), s_->z, n_coef);
}

// P per element and quadrature point as 3x3 matrices (FEAT10Data.cu:613-625).
void TetElementData::RetrievePFromFToCPU(std::vector<std::vector<MatrixXR>>& p_from_F) {
    const size_t n = (size_t)n_elem * nq_ * 9;
    double* d = dalloc<double>(n);
    lib_check(tlfea_qp_to_ref(nn_, n_elem, s_->P_t, d, nullptr), "RetrievePFromFToCPU");
    std::vector<double> h(n);
    d2h(h.data(), d, n);
    dfree(d);
    p_from_F.assign(n_elem, std::vector<MatrixXR>(nq_, MatrixXR(3, 3)));
    for (int e = 0; e < n_elem; e++)
        for (int q = 0; q < nq_; q++) std::copy(&h[((size_t)e * nq_ + q) * 9], &h[((size_t)e * nq_ + q) * 9] + 9, p_from_F[e][q].data());
}

void TetElementData::RetrieveDnDuPreToCPU(std::vector<std::vector<MatrixXR>>& dn_du_pre) {
    const size_t n = (size_t)n_elem * nq_ * nn_ * 3;
    double *d = dalloc<double>(n), *dj = dalloc<double>((size_t)n_elem * nq_);
    lib_check(tlfea_grad_to_ref(nn_, n_elem, s_->grad_t, s_->detJ_t, d, dj, nullptr), "RetrieveDnDuPreToCPU");
    std::vector<double> h(n);
    d2h(h.data(), d, n);
    dfree(d);
    dfree(dj);
    dn_du_pre.assign(n_elem, std::vector<MatrixXR>(nq_, MatrixXR(nn_, 3)));
    for (int e = 0; e < n_elem; e++)
        for (int q = 0; q < nq_; q++)
            std::copy(&h[((size_t)e * nq_ + q) * nn_ * 3], &h[((size_t)e * nq_ + q) * nn_ * 3] + nn_ * 3,
                      dn_du_pre[e][q].data());
}

void TetElementData::RetrieveDetJToCPU(std::vector<std::vector<Real>>& detJ) {
    const size_t n = (size_t)n_elem * nq_;
    double *d = dalloc<double>((size_t)n_elem * nq_ * nn_ * 3), *dj = dalloc<double>(n);
    lib_check(tlfea_grad_to_ref(nn_, n_elem, s_->grad_t, s_->detJ_t, d, dj, nullptr), "RetrieveDetJToCPU");
    std::vector<double> h(n);
    d2h(h.data(), dj, n);
    dfree(d);
    dfree(dj);
    detJ.assign(n_elem, std::vector<Real>(nq_));
    for (int e = 0; e < n_elem; e++)
        for (int q = 0; q < nq_; q++) detJ[e][q] = h[(size_t)e * nq_ + q];
}

void TetElementData::RetrieveConnectivityToCPU(MatrixXi& connectivity) {
    connectivity.resize(n_elem, nn_);
    d2h(connectivity.data(), s_->conn, (size_t)n_elem * nn_);
}

// Legacy ASCII VTK, the layout of FEAT10Data.cu:759-794 (cell type 24 TET10, 10 TET4).
void TetElementData::WriteOutputVTK(const std::string& filename) {
    VectorXR x, y, z;
    RetrievePositionToCPU(x, y, z);
    MatrixXi c;
    RetrieveConnectivityToCPU(c);
    std::ofstream out(filename);
    out << "# vtk DataFile Version 3.0\n";
    out << "T" << nn_ << " mesh output\n";
    out << "ASCII\n";
    out << "DATASET UNSTRUCTURED_GRID\n";
    out << "POINTS " << x.size() << " float\n";
    for (std::ptrdiff_t i = 0; i < x.size(); ++i) out << x(i) << " " << y(i) << " " << z(i) << "\n";
    out << "CELLS " << c.rows() << " " << c.rows() * (nn_ + 1) << "\n";
    for (std::ptrdiff_t i = 0; i < c.rows(); ++i) {
        out << nn_ << " ";
        for (int j = 0; j < nn_; ++j) out << c(i, j) << " ";
        out << "\n";
    }
    out << "CELL_TYPES " << c.rows() << "\n";
    for (std::ptrdiff_t i = 0; i < c.rows(); ++i) out << (nn_ == 10 ? 24 : 10) << "\n";
}

const Real* TetElementData::GetX12DevicePtr() const { return s_->x; }
const Real* TetElementData::GetY12DevicePtr() const { return s_->y; }
const Real* TetElementData::GetZ12DevicePtr() const { return s_->z; }
Real* TetElementData::GetExternalForceDevicePtr() { return s_->f_ext; }
const Real* TetElementData::GetExternalForceDevicePtr() const { return s_->f_ext; }
Real* TetElementData::Get_Constraint_Ptr() { return s_->constraint; }
bool TetElementData::Get_Is_Constraint_Setup() { return is_constraints_setup; }

const tlfea_pattern* TetElementData::Pattern() {
    if (!is_csr_setup) BuildMassCSRPattern();
    return s_->pattern;
}

tlfea_body TetElementData::BodyDescriptor() {
    TetState& s = *s_;
    if (!is_csr_setup) BuildMassCSRPattern();
    if (!s.plan_tried) {
        s.plan_tried = true;
        const int rc = tlfea_force_plan_create(nn_, n_elem, n_coef, s.conn, &s.plan, nullptr);
        if (rc != TLFEA_OK) s.plan = nullptr;  // too scattered for the LDS budget: atomic kernel
    }
    tlfea_body b{};
    b.nn = nn_;
    b.n_elem = n_elem;
    b.n_nodes = n_coef;
    b.conn = s.conn;
    b.x = s.x;
    b.y = s.y;
    b.z = s.z;
    b.grad_t = s.grad_t;
    b.detJ_t = s.detJ_t;
    b.quad = s.quad;
    b.mat = s.mat;
    b.pattern = s.pattern;
    b.m_sell = s.m_sell;
    b.f_ext = s.f_ext;
    b.n_fixed = is_constraints_setup ? n_constraint / 3 : 0;
    b.fixed_nodes = s.fixed;
    b.xref = s.xj;
    b.yref = s.yj;
    b.zref = s.zj;
    b.f_int = s.f_int;
    b.constraint = s.constraint;
    b.force_plan = s.plan;
    // gradients rebuilt from the CalcDnDuPre positions when the plan's blocks fit the LDS
    // budget of that kernel (9 arrays of doubles per block node, 12 with damping)
    if (s.plan && s.have_rest) {
        int block_elems = 0, n_blocks = 0, max_u = 0;
        int64_t n_partials = 0;
        const bool damp = s.mat.eta_damp != 0.0 || s.mat.lambda_damp != 0.0;
        if (tlfea_force_plan_sizes(s.plan, &block_elems, &n_blocks, &n_partials, &max_u) == TLFEA_OK &&
            (size_t)max_u * (damp ? 12 : 9) * sizeof(double) <= 160 * 1024) {
            b.X0 = s.x0;
            b.Y0 = s.y0;
            b.Z0 = s.z0;
        }
    }
    return b;
}

// FEAT10Data.cuh:659-723
void TetElementData::Destroy() {
    TetState& s = *s_;
    for (double** p : {&s.x, &s.y, &s.z, &s.xj, &s.yj, &s.zj, &s.x0, &s.y0, &s.z0, &s.grad_t, &s.detJ_t, &s.F_t, &s.P_t, &s.Fd_t, &s.Pv_t,
                       &s.f_int, &s.f_ext, &s.constraint, &s.m_sell, &s.j_val, &s.jt_val})
        dfree(*p);
    for (int** p : {&s.conn, &s.fixed, &s.j_off, &s.j_col, &s.jt_off, &s.jt_col}) dfree(*p);
    if (s.pattern) tlfea_pattern_destroy(s.pattern);
    if (s.plan) tlfea_force_plan_destroy(s.plan);
    s.pattern = nullptr;
    s.plan = nullptr;
    s.plan_tried = false;
    s.have_rest = false;
    s.initialized = false;
    is_setup = is_constraints_setup = is_csr_setup = is_cj_csr_setup = is_j_csr_setup = false;
}

}  // namespace detail
}  // namespace tlfea
