// FP64 CPU oracle for the TLFEA solve path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the C++ drop-in
// classes, the Python host package) links, loads or calls this file; only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, to check results against.
//
// It restates, in plain sequential C++ (OpenMP over elements where the reference's
// arithmetic is per-element independent), the arithmetic of the reference files below;
// each function cites the lines it follows.  It adds what the reference does not have:
// brute-force CSR patterns (std::set per row, independent of the sort/unique route),
// a banded Cholesky direct solve, and residual helpers.
//
// Array conventions are the reference's (`src/elements/FEAT10Data.cuh:41-51,113-161`):
//   conn   int32 column-major   conn[e + a*n_elem]
//   x,y,z  float64 SoA          x[node]
//   grad   float64              grad[(e*nq + q)*nn*3 + a + nn*d]  (dN_a/dX_d)
//   detJ   float64              detJ[e*nq + q]
//   F,P    float64              F[(e*nq + q)*9 + i + 3*j]          (column-major 3x3)
//   f      float64 AoS          f[3*node + d]
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

const int EDGES[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

// Gaussian elimination with partial pivoting on [A|b]; a pivot below 1e-14 yields x = 0.
// Follows `src/elements/FEAT10DataFunc.cuh:32-84`.
void solve3(const double A[3][3], const double b[3], double x[3]) {
    double aug[3][4];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) aug[i][j] = A[i][j];
        aug[i][3] = b[i];
    }
    for (int k = 0; k < 3; k++) {
        int p = k;
        double mv = std::fabs(aug[k][k]);
        for (int i = k + 1; i < 3; i++)
            if (std::fabs(aug[i][k]) > mv) { mv = std::fabs(aug[i][k]); p = i; }
        if (p != k)
            for (int j = 0; j < 4; j++) std::swap(aug[k][j], aug[p][j]);
        if (std::fabs(aug[k][k]) < 1e-14) { x[0] = x[1] = x[2] = 0.0; return; }
        for (int i = k + 1; i < 3; i++) {
            double fct = aug[i][k] / aug[k][k];
            for (int j = k; j < 4; j++) aug[i][j] -= fct * aug[k][j];
        }
    }
    x[2] = aug[2][3] / aug[2][2];
    x[1] = (aug[1][3] - aug[1][2] * x[2]) / aug[1][1];
    x[0] = (aug[0][3] - aug[0][2] * x[2] - aug[0][1] * x[1]) / aug[0][0];
}

inline double det3(const double J[3][3]) {
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// dN/dxi of the quadratic tet at barycentric (1-xi-eta-zeta, xi, eta, zeta).
// `src/elements/FEAT10Data.cu:104-136`.
void t10_dN_dxi(double xi, double eta, double zeta, double dN[10][3]) {
    double L[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
    const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 4; i++) {
        double f = 4.0 * L[i] - 1.0;
        for (int j = 0; j < 3; j++) dN[i][j] = f * dL[i][j];
    }
    for (int k = 0; k < 6; k++) {
        int i = EDGES[k][0], j = EDGES[k][1];
        for (int d = 0; d < 3; d++) dN[k + 4][d] = 4.0 * (L[i] * dL[j][d] + L[j] * dL[i][d]);
    }
}

// Quadratic tet shape functions N (`src/elements/FEAT10Data.cu:229-243`).
void t10_N(double xi, double eta, double zeta, double N[10]) {
    double L[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
    for (int k = 0; k < 4; k++) N[k] = L[k] * (2.0 * L[k] - 1.0);
    for (int k = 0; k < 6; k++) N[k + 4] = 4.0 * L[EDGES[k][0]] * L[EDGES[k][1]];
}

// SVK first Piola stress: P = lambda (tr(F^T F)/2 - 3/2) F + mu (F F^T F - F).
// `src/materials/SVK.cuh:17-39` with the F^T F, F F^T, F F^T F products of
// `src/elements/FEAT10DataFunc.cuh:236-278`.
void svk_P(const double F[3][3], double lambda, double mu, double P[3][3]) {
    double FtF[3][3] = {{0}}, FFt[3][3] = {{0}}, FFtF[3][3] = {{0}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) FtF[i][j] += F[k][i] * F[k][j];
    double tr = FtF[0][0] + FtF[1][1] + FtF[2][2];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) FFt[i][j] += F[i][k] * F[j][k];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) FFtF[i][j] += FFt[i][k] * F[k][j];
    double lf = lambda * (0.5 * tr - 1.5);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) P[i][j] = lf * F[i][j] + mu * (FFtF[i][j] - F[i][j]);
}

// Kelvin-Voigt viscous Piola: Edot = (Fdot^T F + F^T Fdot)/2,
// S = 2 eta Edot + lambda_d tr(Edot) I, P_vis = F S.  `FEAT10DataFunc.cuh:135-218`.
void kv_Pvis(const double F[3][3], const double Fd[3][3], double eta, double ld, double Pv[3][3]) {
    double A[3][3] = {{0}}, B[3][3] = {{0}}, Ed[3][3], S[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                A[i][j] += Fd[k][i] * F[k][j];
                B[i][j] += F[k][i] * Fd[k][j];
            }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ed[i][j] = 0.5 * (A[i][j] + B[i][j]);
    double tr = Ed[0][0] + Ed[1][1] + Ed[2][2];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i][j] = 2.0 * eta * Ed[i][j] + ld * tr * (i == j ? 1.0 : 0.0);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Pv[i][j] = 0.0;
            for (int k = 0; k < 3; k++) Pv[i][j] += F[i][k] * S[k][j];
        }
}

// Compressible Mooney-Rivlin (`src/materials/MooneyRivlin.cuh:45-227`):
//   W = mu10 (J^-2/3 I1 - 3) + mu01 (J^-4/3 I2 - 3) + kappa/2 (J - 1)^2,
//   I1 = tr C, I2 = (I1^2 - tr C^2)/2, C = F^T F, J = det F,
//   P = dW/dF = t1 (F - I1/3 F^-T) + t2 (I1 F - F C - 2 I2/3 F^-T) + t3 F^-T,
//   t1 = 2 mu10 J^-2/3, t2 = 2 mu01 J^-4/3, t3 = kappa (J - 1) J.
// F^-T divides by det F clamped to +-1e-12 in magnitude, as the reference does
// (MooneyRivlin.cuh:24-43).
struct MrParts {
    double C[3][3], FC[3][3], FFT[3][3], FiT[3][3], I1, I2, J, t1, t2, t3;
};

void mr_parts(const double F[3][3], double mu10, double mu01, double kappa, MrParts& m) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            m.C[i][j] = m.FFT[i][j] = 0.0;
            for (int k = 0; k < 3; k++) {
                m.C[i][j] += F[k][i] * F[k][j];
                m.FFT[i][j] += F[i][k] * F[j][k];
            }
        }
    m.I1 = m.C[0][0] + m.C[1][1] + m.C[2][2];
    double trC2 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) trC2 += m.C[i][k] * m.C[k][i];
    m.I2 = 0.5 * (m.I1 * m.I1 - trC2);
    m.J = det3(F);
    double dsafe = m.J;
    if (std::fabs(dsafe) < 1e-12) dsafe = dsafe >= 0.0 ? 1e-12 : -1e-12;
    // F^-T_ij = cof(F)_ij / det F, cof(F)_ij = (-1)^(i+j) minor_ij
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            m.FiT[i][j] = (F[i1][j1] * F[i2][j2] - F[i1][j2] * F[i2][j1]) / dsafe;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            m.FC[i][j] = 0.0;
            for (int k = 0; k < 3; k++) m.FC[i][j] += F[i][k] * m.C[k][j];
        }
    const double c = std::cbrt(m.J);
    m.t1 = 2.0 * mu10 / (c * c);
    m.t2 = 2.0 * mu01 / (c * c * c * c);
    m.t3 = kappa * (m.J - 1.0) * m.J;
}

void mr_P(const double F[3][3], double mu10, double mu01, double kappa, double P[3][3]) {
    MrParts m;
    mr_parts(F, mu10, mu01, kappa, m);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            P[i][j] = m.t1 * (F[i][j] - m.I1 / 3.0 * m.FiT[i][j]) +
                      m.t2 * (m.I1 * F[i][j] - m.FC[i][j] - 2.0 * m.I2 / 3.0 * m.FiT[i][j]) + m.t3 * m.FiT[i][j];
}

// A_ijkl = dP_ij/dF_kl (`src/materials/MooneyRivlin.cuh:106-227`), from
//   dJ/dF_kl = J F^-T_kl, d(F^-T)_ij/dF_kl = -F^-T_il F^-T_kj, dI1/dF_kl = 2 F_kl,
//   dI2/dF_kl = 2 (I1 F - F C)_kl, d(F C)_ij/dF_kl = d_ik C_lj + F_il F_kj + d_jl (F F^T)_ik.
void mr_A(const double F[3][3], double mu10, double mu01, double kappa, double A[3][3][3][3]) {
    MrParts m;
    mr_parts(F, mu10, mu01, kappa, m);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double T1 = F[i][j] - m.I1 / 3.0 * m.FiT[i][j];
            const double T2 = m.I1 * F[i][j] - m.FC[i][j] - 2.0 * m.I2 / 3.0 * m.FiT[i][j];
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) {
                    const double dik = i == k, djl = j == l;
                    const double dJ = m.J * m.FiT[k][l];  // dJ/dF_kl
                    const double dFiT = -m.FiT[i][l] * m.FiT[k][j];
                    const double dI1 = 2.0 * F[k][l];
                    const double dI2 = 2.0 * (m.I1 * F[k][l] - m.FC[k][l]);
                    const double dFC = dik * m.C[l][j] + F[i][l] * F[k][j] + djl * m.FFT[i][k];
                    // t1 ~ J^-2/3, t2 ~ J^-4/3: dt = t * (-p/3) dJ / J
                    const double dt1 = -2.0 / 3.0 * m.t1 * m.FiT[k][l];
                    const double dt2 = -4.0 / 3.0 * m.t2 * m.FiT[k][l];
                    const double dt3 = kappa * (2.0 * m.J - 1.0) * dJ;
                    const double dT1 = dik * djl - dI1 / 3.0 * m.FiT[i][j] - m.I1 / 3.0 * dFiT;
                    const double dT2 = dI1 * F[i][j] + m.I1 * dik * djl - dFC - 2.0 / 3.0 * dI2 * m.FiT[i][j] -
                                       2.0 * m.I2 / 3.0 * dFiT;
                    A[i][j][k][l] = dt1 * T1 + m.t1 * dT1 + dt2 * T2 + m.t2 * dT2 + dt3 * m.FiT[i][j] + m.t3 * dFiT;
                }
        }
}

double mr_W(const double F[3][3], double mu10, double mu01, double kappa) {
    MrParts m;
    mr_parts(F, mu10, mu01, kappa, m);
    const double c = std::cbrt(m.J);
    return mu10 * (m.I1 / (c * c) - 3.0) + mu01 * (m.I2 / (c * c * c * c) - 3.0) +
           0.5 * kappa * (m.J - 1.0) * (m.J - 1.0);
}

// Material of every call below: 0 SVK (lambda, mu arguments), 1 Mooney-Rivlin with these
// coefficients (`d_material_model`, FEAT10Data.cuh:533-587).  Set by orc_set_material.
int g_model = 0;
double g_mu10 = 0.0, g_mu01 = 0.0, g_kappa = 0.0;

void material_P(const double F[3][3], double lambda, double mu, double P[3][3]) {
    if (g_model == 1)
        mr_P(F, g_mu10, g_mu01, g_kappa, P);
    else
        svk_P(F, lambda, mu, P);
}

struct Elem {
    int n_elem, nn, nq;
    const int* conn;
    int node(int e, int a) const { return conn[e + (int64_t)a * n_elem]; }
};

}  // namespace

// ---------------------------------------------------------------------------------------
// Quadrature (`src/utils/quadrature_utils.h:125-179`)
// ---------------------------------------------------------------------------------------
ORC_API void orc_quad_t10(double* qx, double* qy, double* qz, double* qw) {
    const double bary[5][4] = {{0.25, 0.25, 0.25, 0.25},
                               {0.5, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                               {1.0 / 6.0, 0.5, 1.0 / 6.0, 1.0 / 6.0},
                               {1.0 / 6.0, 1.0 / 6.0, 0.5, 1.0 / 6.0},
                               {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.5}};
    const double w[5] = {-4.0 / 5.0, 9.0 / 20.0, 9.0 / 20.0, 9.0 / 20.0, 9.0 / 20.0};
    for (int q = 0; q < 5; q++) {
        qx[q] = bary[q][1];
        qy[q] = bary[q][2];
        qz[q] = bary[q][3];
        qw[q] = w[q] * (1.0 / 6.0);
    }
}

ORC_API void orc_quad_t4(double* qx, double* qy, double* qz, double* qw) {
    qx[0] = qy[0] = qz[0] = 0.25;
    qw[0] = 1.0 / 6.0;
}

// ---------------------------------------------------------------------------------------
// Reference shape-function gradients and det J
// TET10: `src/elements/FEAT10Data.cu:92-198` (partial-pivot solve of J^T g_a = dN_a/dxi).
// TET4:  `src/elements/FEAT4Data.cu:98-190` (explicit cofactor inverse, 1/det or 0).
// ---------------------------------------------------------------------------------------
ORC_API void orc_t10_grad(int n_elem, const int* conn, const double* x, const double* y, const double* z,
                          const double* qx, const double* qy, const double* qz, double* grad, double* detJ) {
    Elem E{n_elem, 10, 5, conn};
#pragma omp parallel for schedule(static)
    for (int e = 0; e < n_elem; e++) {
        double X[10][3];
        for (int a = 0; a < 10; a++) {
            int g = E.node(e, a);
            X[a][0] = x[g]; X[a][1] = y[g]; X[a][2] = z[g];
        }
        for (int q = 0; q < 5; q++) {
            double dN[10][3];
            t10_dN_dxi(qx[q], qy[q], qz[q], dN);
            double J[3][3] = {{0}};
            for (int a = 0; a < 10; a++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) J[i][j] += X[a][i] * dN[a][j];
            detJ[e * 5 + q] = det3(J);
            double JT[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) JT[i][j] = J[j][i];
            double* G = grad + ((int64_t)e * 5 + q) * 30;
            for (int a = 0; a < 10; a++) {
                double g[3];
                solve3(JT, dN[a], g);
                for (int d = 0; d < 3; d++) G[a + 10 * d] = g[d];
            }
        }
    }
}

ORC_API void orc_t4_grad(int n_elem, const int* conn, const double* x, const double* y, const double* z,
                         double* grad, double* detJ) {
    Elem E{n_elem, 4, 1, conn};
    const double dN[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma omp parallel for schedule(static)
    for (int e = 0; e < n_elem; e++) {
        double X[4][3];
        for (int a = 0; a < 4; a++) {
            int g = E.node(e, a);
            X[a][0] = x[g]; X[a][1] = y[g]; X[a][2] = z[g];
        }
        double J[3][3] = {{0}};
        for (int a = 0; a < 4; a++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) J[i][j] += X[a][i] * dN[a][j];
        detJ[e] = det3(J);
        double T[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) T[i][j] = J[j][i];
        double dt = det3(T);
        double id = (dt != 0.0) ? 1.0 / dt : 0.0;
        double Ti[3][3];
        Ti[0][0] = (T[1][1] * T[2][2] - T[1][2] * T[2][1]) * id;
        Ti[0][1] = -(T[0][1] * T[2][2] - T[0][2] * T[2][1]) * id;
        Ti[0][2] = (T[0][1] * T[1][2] - T[0][2] * T[1][1]) * id;
        Ti[1][0] = -(T[1][0] * T[2][2] - T[1][2] * T[2][0]) * id;
        Ti[1][1] = (T[0][0] * T[2][2] - T[0][2] * T[2][0]) * id;
        Ti[1][2] = -(T[0][0] * T[1][2] - T[0][2] * T[1][0]) * id;
        Ti[2][0] = (T[1][0] * T[2][1] - T[1][1] * T[2][0]) * id;
        Ti[2][1] = -(T[0][0] * T[2][1] - T[0][1] * T[2][0]) * id;
        Ti[2][2] = (T[0][0] * T[1][1] - T[0][1] * T[1][0]) * id;
        double* G = grad + (int64_t)e * 12;
        for (int a = 0; a < 4; a++)
            for (int i = 0; i < 3; i++) {
                double s = 0.0;
                for (int j = 0; j < 3; j++) s += Ti[i][j] * dN[a][j];
                G[a + 4 * i] = s;
            }
    }
}

// ---------------------------------------------------------------------------------------
// F and P at every (e,q).  `src/elements/FEAT10DataFunc.cuh:86-291` (SVK branch);
// TET4 `src/elements/FEAT4DataFunc.cuh:33-200`.  v == nullptr or eta == lambda_d == 0
// means no damping (Fdot = P_vis = 0 stored).  Any of F/P/Fdot/Pvis may be null.
// ---------------------------------------------------------------------------------------
ORC_API void orc_calc_p(int nn, int nq, int n_elem, const int* conn, const double* x, const double* y,
                        const double* z, const double* grad, double lambda, double mu, double eta, double lambda_d,
                        const double* v, double* Fo, double* Po, double* Fdo, double* Pvo) {
    Elem E{n_elem, nn, nq, conn};
    const bool damp = (v != nullptr) && (eta != 0.0 || lambda_d != 0.0);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < n_elem; e++) {
        for (int q = 0; q < nq; q++) {
            const double* G = grad + ((int64_t)e * nq + q) * nn * 3;
            double F[3][3] = {{0}}, Fd[3][3] = {{0}}, P[3][3], Pv[3][3] = {{0}};
            for (int a = 0; a < nn; a++) {
                int g = E.node(e, a);
                double xa[3] = {x[g], y[g], z[g]};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) F[i][j] += xa[i] * G[a + nn * j];
            }
            if (damp) {
                for (int a = 0; a < nn; a++) {
                    int g = E.node(e, a);
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) Fd[i][j] += v[3 * g + i] * G[a + nn * j];
                }
                kv_Pvis(F, Fd, eta, lambda_d, Pv);
            }
            material_P(F, lambda, mu, P);
            int64_t o = ((int64_t)e * nq + q) * 9;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    if (Fo) Fo[o + i + 3 * j] = F[i][j];
                    if (Po) Po[o + i + 3 * j] = P[i][j] + Pv[i][j];
                    if (Fdo) Fdo[o + i + 3 * j] = Fd[i][j];
                    if (Pvo) Pvo[o + i + 3 * j] = Pv[i][j];
                }
        }
    }
}

// f_int[3*node+d] += sum_q P_q g_a(q) detJ_q w_q for every (e,a).  Accumulates, as the
// reference's standalone CalcInternalForce does.  `src/elements/FEAT10DataFunc.cuh:404-460`.
ORC_API void orc_fint_from_p(int nn, int nq, int n_elem, const int* conn, const double* P, const double* grad,
                             const double* detJ, const double* qw, double* f) {
    Elem E{n_elem, nn, nq, conn};
    for (int e = 0; e < n_elem; e++) {
        for (int a = 0; a < nn; a++) {
            double fa[3] = {0, 0, 0};
            for (int q = 0; q < nq; q++) {
                const double* Pq = P + ((int64_t)e * nq + q) * 9;
                const double* G = grad + ((int64_t)e * nq + q) * nn * 3;
                double g[3] = {G[a], G[a + nn], G[a + 2 * nn]};
                double dV = detJ[e * nq + q] * qw[q];
                for (int i = 0; i < 3; i++) {
                    double c = 0.0;
                    for (int j = 0; j < 3; j++) c += Pq[i + 3 * j] * g[j];
                    fa[i] += c * dV;
                }
            }
            int n = E.node(e, a);
            for (int i = 0; i < 3; i++) f[3 * n + i] += fa[i];
        }
    }
}

// Convenience: f_int = 0 then P and f_int at positions (x,y,z), velocity v for damping.
ORC_API void orc_internal_force(int nn, int nq, int n_elem, int n_nodes, const int* conn, const double* x,
                                const double* y, const double* z, const double* grad, const double* detJ,
                                const double* qw, double lambda, double mu, double eta, double lambda_d,
                                const double* v, double* f) {
    std::vector<double> P((size_t)n_elem * nq * 9);
    orc_calc_p(nn, nq, n_elem, conn, x, y, z, grad, lambda, mu, eta, lambda_d, v, nullptr, P.data(), nullptr,
               nullptr);
    std::memset(f, 0, sizeof(double) * 3 * (size_t)n_nodes);
    orc_fint_from_p(nn, nq, n_elem, conn, P.data(), grad, detJ, qw, f);
}

// Total strain energy  sum_e,q W(F) detJ w  with W = lambda/2 (tr E)^2 + mu tr(E^2) (SVK)
// or the Mooney-Rivlin W above.
// (Not in the reference: its gradient w.r.t. x is f_int, used to pin f_int by finite
// differences.)
ORC_API double orc_energy(int nn, int nq, int n_elem, const int* conn, const double* x, const double* y,
                          const double* z, const double* grad, const double* detJ, const double* qw, double lambda,
                          double mu) {
    Elem E{n_elem, nn, nq, conn};
    double tot = 0.0;
    for (int e = 0; e < n_elem; e++)
        for (int q = 0; q < nq; q++) {
            const double* G = grad + ((int64_t)e * nq + q) * nn * 3;
            double F[3][3] = {{0}};
            for (int a = 0; a < nn; a++) {
                int g = E.node(e, a);
                double xa[3] = {x[g], y[g], z[g]};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) F[i][j] += xa[i] * G[a + nn * j];
            }
            double Em[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double c = 0;
                    for (int k = 0; k < 3; k++) c += F[k][i] * F[k][j];
                    Em[i][j] = 0.5 * (c - (i == j ? 1.0 : 0.0));
                }
            double tr = Em[0][0] + Em[1][1] + Em[2][2];
            double trE2 = 0;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) trE2 += Em[i][j] * Em[j][i];
            const double W = g_model == 1 ? mr_W(F, g_mu10, g_mu01, g_kappa) : 0.5 * lambda * tr * tr + mu * trE2;
            tot += W * detJ[e * nq + q] * qw[q];
        }
    return tot;
}

// ---------------------------------------------------------------------------------------
// CSR patterns, by brute force (a std::set per row).  The reference builds the same
// patterns by sorting (row<<32|col) keys: node level `src/elements/FEAT10Data.cu:34-74,
// 361-417`; DOF level `src/solvers/LinearStaticSolver.cu:75-122,240-294`.  Rows list every
// node (DOF) sharing an element with the row node (DOF), diagonal included, ascending.
// Two passes: call with cols == nullptr to get nnz (returned), then with storage.
// ---------------------------------------------------------------------------------------
ORC_API int64_t orc_node_pattern(int nn, int n_elem, int n_nodes, const int* conn, int* offsets, int* cols) {
    std::vector<std::set<int>> rows(n_nodes);
    for (int e = 0; e < n_elem; e++)
        for (int a = 0; a < nn; a++)
            for (int b = 0; b < nn; b++) rows[conn[e + (int64_t)a * n_elem]].insert(conn[e + (int64_t)b * n_elem]);
    int64_t nnz = 0;
    for (int i = 0; i < n_nodes; i++) {
        if (offsets) offsets[i] = (int)nnz;
        if (cols)
            for (int c : rows[i]) cols[nnz++] = c;
        else
            nnz += (int64_t)rows[i].size();
    }
    if (offsets) offsets[n_nodes] = (int)nnz;
    return nnz;
}

ORC_API int64_t orc_dof_pattern(int nn, int n_elem, int n_nodes, const int* conn, int* offsets, int* cols) {
    int n_dof = 3 * n_nodes;
    std::vector<std::set<int>> rows(n_dof);
    for (int e = 0; e < n_elem; e++)
        for (int a = 0; a < nn; a++)
            for (int b = 0; b < nn; b++) {
                int i = conn[e + (int64_t)a * n_elem], j = conn[e + (int64_t)b * n_elem];
                for (int d = 0; d < 3; d++)
                    for (int c = 0; c < 3; c++) rows[3 * i + d].insert(3 * j + c);
            }
    int64_t nnz = 0;
    for (int r = 0; r < n_dof; r++) {
        if (offsets) offsets[r] = (int)nnz;
        if (cols)
            for (int c : rows[r]) cols[nnz++] = c;
        else
            nnz += (int64_t)rows[r].size();
    }
    if (offsets) offsets[n_dof] = (int)nnz;
    return nnz;
}

static int find_col(const int* cols, int lo, int hi, int target) {
    const int* p = std::lower_bound(cols + lo, cols + hi, target);
    return (p != cols + hi && *p == target) ? (int)(p - cols) : -1;
}

// Consistent mass on the node-level pattern.  TET10: M_ij += sum_q rho N_i N_j detJ_q w_q
// with the 5-point rule (`src/elements/FEAT10Data.cu:200-271`); TET4: 1-point centroid rule,
// every M_ij = rho detJ/96 (`src/elements/FEAT4Data.cu:197-243`).  Values are zeroed first.
ORC_API void orc_mass_values(int nn, int nq, int n_elem, const int* conn, const double* detJ, const double* qx,
                             const double* qy, const double* qz, const double* qw, double rho, const int* offsets,
                             const int* cols, double* vals, int n_nodes) {
    std::memset(vals, 0, sizeof(double) * (size_t)offsets[n_nodes]);
    for (int e = 0; e < n_elem; e++)
        for (int i = 0; i < nn; i++)
            for (int j = 0; j < nn; j++) {
                double m = 0.0;
                for (int q = 0; q < nq; q++) {
                    double N[10];
                    if (nn == 10) {
                        t10_N(qx[q], qy[q], qz[q], N);
                    } else {
                        N[0] = 1.0 - qx[q] - qy[q] - qz[q]; N[1] = qx[q]; N[2] = qy[q]; N[3] = qz[q];
                    }
                    m += rho * N[i] * N[j] * detJ[e * nq + q] * qw[q];
                }
                int gi = conn[e + (int64_t)i * n_elem], gj = conn[e + (int64_t)j * n_elem];
                int k = find_col(cols, offsets[gi], offsets[gi + 1], gj);
                if (k >= 0) vals[k] += m;
            }
}

// ---------------------------------------------------------------------------------------
// Tangent stiffness on the DOF-level pattern at the current positions, scaled by h, plus
// the unscaled Kelvin-Voigt block when damping is on.
// `src/elements/FEAT10DataFunc.cuh:512-768`, `src/materials/SVK.cuh:41-67`; the
// Mooney-Rivlin block from mr_A.
// ---------------------------------------------------------------------------------------
// Per (e, q) kinematics of the tangent: node ids, g_a, F g_a, tr E, F F^T, dV and (Mooney-
// Rivlin) A = dP/dF.
struct TanPoint {
    int gn[10];
    double g[10][3], Fh[10][3], FFT[3][3], trE, dV;
    double A[3][3][3][3];
};

static void tangent_point(const Elem& E, int e, int q, const double* x, const double* y, const double* z,
                          const double* grad, const double* detJ, const double* qw, TanPoint& t) {
    const int nn = E.nn, nq = E.nq;
    const double* G = grad + ((int64_t)e * nq + q) * nn * 3;
    double F[3][3] = {{0}};
    for (int a = 0; a < nn; a++) {
        t.gn[a] = E.node(e, a);
        double xa[3] = {x[t.gn[a]], y[t.gn[a]], z[t.gn[a]]};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) F[i][j] += xa[i] * G[a + nn * j];
    }
    double C[3][3] = {{0}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            t.FFT[i][j] = 0.0;
            for (int k = 0; k < 3; k++) {
                C[i][j] += F[k][i] * F[k][j];
                t.FFT[i][j] += F[i][k] * F[j][k];
            }
        }
    t.trE = 0.5 * (C[0][0] + C[1][1] + C[2][2] - 3.0);
    for (int a = 0; a < nn; a++) {
        for (int d = 0; d < 3; d++) t.g[a][d] = G[a + nn * d];
        for (int r = 0; r < 3; r++) {
            t.Fh[a][r] = 0.0;
            for (int c = 0; c < 3; c++) t.Fh[a][r] += F[r][c] * t.g[a][c];
        }
    }
    t.dV = detJ[e * nq + q] * qw[q];
    // Mooney-Rivlin block: K_ij[d][c] = sum_JL A[d][J][c][L] g_i[J] g_j[L] dV
    // (FEAT10DataFunc.cuh:607-640)
    if (g_model == 1) mr_A(F, g_mu10, g_mu01, g_kappa, t.A);
}

// Entry [d][c] of the (i, j) node block: h K + C (Kelvin-Voigt, unscaled) at one point.
// K (SVK.cuh:41-67): lambda Fh_i,d Fh_j,c + lambda trE (g_i.g_j) delta + mu (Fh_j.Fh_i) delta
// + mu Fh_j,d Fh_i,c + mu (g_i.g_j) (F F^T)_dc - mu (g_i.g_j) delta, times dV.
// C (FEAT10DataFunc.cuh:686-767): eta Fh_j (x) Fh_i + eta F F^T (g_i.g_j) + lambda_d Fh_i (x) Fh_j.
static inline double tangent_entry(const TanPoint& t, int i, int j, int d, int c, double lambda, double mu,
                                   double eta_d, double lambda_d, double h) {
    const double hij = t.g[j][0] * t.g[i][0] + t.g[j][1] * t.g[i][1] + t.g[j][2] * t.g[i][2];
    const double FjFi = t.Fh[j][0] * t.Fh[i][0] + t.Fh[j][1] * t.Fh[i][1] + t.Fh[j][2] * t.Fh[i][2];
    const double dl = (d == c) ? 1.0 : 0.0;
    double K = (lambda * t.Fh[i][d] * t.Fh[j][c] + lambda * t.trE * hij * dl + mu * FjFi * dl +
                mu * t.Fh[j][d] * t.Fh[i][c] + mu * hij * t.FFT[d][c] - mu * hij * dl) *
               t.dV;
    if (g_model == 1) {
        K = 0.0;
        for (int J = 0; J < 3; J++)
            for (int L = 0; L < 3; L++) K += t.A[d][J][c][L] * t.g[i][J] * t.g[j][L];
        K *= t.dV;
    }
    double Cv = 0.0;
    if (eta_d != 0.0 || lambda_d != 0.0)
        Cv = (eta_d * (t.Fh[j][d] * t.Fh[i][c]) + eta_d * t.FFT[d][c] * hij + lambda_d * (t.Fh[i][d] * t.Fh[j][c])) *
             t.dV;
    return h * K + Cv;
}

ORC_API void orc_stiffness_values(int nn, int nq, int n_elem, const int* conn, const double* x, const double* y,
                                  const double* z, const double* grad, const double* detJ, const double* qw,
                                  double lambda, double mu, double eta_d, double lambda_d, double h,
                                  const int* offsets, const int* cols, double* vals) {
    Elem E{n_elem, nn, nq, conn};
    static thread_local TanPoint t;
    for (int e = 0; e < n_elem; e++)
        for (int q = 0; q < nq; q++) {
            tangent_point(E, e, q, x, y, z, grad, detJ, qw, t);
            for (int i = 0; i < nn; i++)
                for (int j = 0; j < nn; j++)
                    for (int d = 0; d < 3; d++) {
                        const int row = 3 * t.gn[i] + d;
                        for (int c = 0; c < 3; c++) {
                            const int k = find_col(cols, offsets[row], offsets[row + 1], 3 * t.gn[j] + c);
                            if (k >= 0) vals[k] += tangent_entry(t, i, j, d, c, lambda, mu, eta_d, lambda_d, h);
                        }
                    }
        }
}

// The same assembly with the elements split over the OpenMP threads and atomic adds into
// the values: for timing the CPU baseline on all cores (the order of the adds, so the last
// bits, depend on the thread schedule; the tests use the sequential assembly above).
ORC_API void orc_stiffness_values_omp(int nn, int nq, int n_elem, const int* conn, const double* x, const double* y,
                                      const double* z, const double* grad, const double* detJ, const double* qw,
                                      double lambda, double mu, double eta_d, double lambda_d, double h,
                                      const int* offsets, const int* cols, double* vals) {
    Elem E{n_elem, nn, nq, conn};
#pragma omp parallel
    {
        TanPoint t;
#pragma omp for schedule(static)
        for (int e = 0; e < n_elem; e++)
            for (int q = 0; q < nq; q++) {
                tangent_point(E, e, q, x, y, z, grad, detJ, qw, t);
                for (int i = 0; i < nn; i++)
                    for (int j = 0; j < nn; j++)
                        for (int d = 0; d < 3; d++) {
                            const int row = 3 * t.gn[i] + d;
                            for (int c = 0; c < 3; c++) {
                                const int k = find_col(cols, offsets[row], offsets[row + 1], 3 * t.gn[j] + c);
                                if (k < 0) continue;
                                const double v = tangent_entry(t, i, j, d, c, lambda, mu, eta_d, lambda_d, h);
#pragma omp atomic
                                vals[k] += v;
                            }
                        }
            }
    }
}

// out = (h K(x) + C) p element by element, nothing assembled: the operator the assembled
// values above hold, applied as sum_e sum_q K_e,q p_e.  Elements are split over the OpenMP
// threads, each summing into its own array; the arrays are added in thread order.  (Not in
// the reference: the true residual f - K u of large solutions without the 2 GB CSR.)
ORC_API void orc_tangent_apply(int nn, int nq, int n_elem, int n_nodes, const int* conn, const double* x,
                               const double* y, const double* z, const double* grad, const double* detJ,
                               const double* qw, double lambda, double mu, double eta_d, double lambda_d, double h,
                               const double* p, double* out) {
    Elem E{n_elem, nn, nq, conn};
    const int64_t n3 = 3 * (int64_t)n_nodes;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> part(nt);
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        std::vector<double>& acc = part[tid];
        acc.assign(n3, 0.0);
        TanPoint t;
#pragma omp for schedule(static)
        for (int e = 0; e < n_elem; e++)
            for (int q = 0; q < nq; q++) {
                tangent_point(E, e, q, x, y, z, grad, detJ, qw, t);
                for (int i = 0; i < nn; i++)
                    for (int d = 0; d < 3; d++) {
                        double s = 0.0;
                        for (int j = 0; j < nn; j++)
                            for (int c = 0; c < 3; c++)
                                s += tangent_entry(t, i, j, d, c, lambda, mu, eta_d, lambda_d, h) * p[3 * t.gn[j] + c];
                        acc[3 * t.gn[i] + d] += s;
                    }
            }
    }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n3; k++) {
        double s = 0.0;
        for (int t = 0; t < nt; t++) s += part[t][k];
        out[k] = s;
    }
}

// Dirichlet elimination: fixed rows -> identity row, f = 0; free rows -> fixed columns 0.
// `src/solvers/LinearStaticSolver.cu:147-184,316-342`.
ORC_API void orc_apply_dirichlet(int n_dof, const int* offsets, const int* cols, double* vals, double* f,
                                 int n_fixed_nodes, const int* fixed_nodes) {
    std::vector<char> fixed(n_dof, 0);
    for (int k = 0; k < n_fixed_nodes; k++)
        for (int d = 0; d < 3; d++) fixed[3 * fixed_nodes[k] + d] = 1;
    for (int r = 0; r < n_dof; r++) {
        if (fixed[r]) {
            for (int k = offsets[r]; k < offsets[r + 1]; k++) vals[k] = (cols[k] == r) ? 1.0 : 0.0;
            f[r] = 0.0;
        } else {
            for (int k = offsets[r]; k < offsets[r + 1]; k++)
                if (fixed[cols[k]]) vals[k] = 0.0;
        }
    }
}

ORC_API void orc_spmv(int n, const int* offsets, const int* cols, const double* vals, const double* xv, double* yv) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++) {
        double s = 0.0;
        for (int k = offsets[r]; k < offsets[r + 1]; k++) s += vals[k] * xv[cols[k]];
        yv[r] = s;
    }
}

static double dot(int n, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

// Unpreconditioned CG from u = 0 with the reference's stopping rule sqrt(rr/rr0) < tol,
// abort on p.Kp <= 0.  `src/solvers/LinearStaticSolver.cu:348-443`.
// Returns the iteration count (-1 on abort); *res_out = final relative residual.
ORC_API int orc_cg(int n, const int* offsets, const int* cols, const double* vals, const double* f, double* u,
                   double tol, int max_iter, double* res_out) {
    std::vector<double> r(f, f + n), p(f, f + n), Kp(n);
    std::fill(u, u + n, 0.0);
    double rr0 = dot(n, r.data(), r.data());
    if (rr0 == 0.0) { *res_out = 0.0; return 0; }
    double rr = rr0;
    int iters = max_iter;
    double last = 1.0;
    for (int it = 0; it < max_iter; it++) {
        orc_spmv(n, offsets, cols, vals, p.data(), Kp.data());
        double pKp = dot(n, p.data(), Kp.data());
        if (pKp <= 0.0) { *res_out = last; return -1; }
        double alpha = rr / pKp;
        for (int i = 0; i < n; i++) u[i] += alpha * p[i];
        for (int i = 0; i < n; i++) r[i] -= alpha * Kp[i];
        double rn = dot(n, r.data(), r.data());
        double rel = std::sqrt(rn / rr0);
        last = rel;
        if (rel < tol) { iters = it + 1; break; }
        double beta = rn / rr;
        for (int i = 0; i < n; i++) p[i] = beta * p[i] + r[i];
        rr = rn;
    }
    *res_out = last;
    return iters;
}

// ---------------------------------------------------------------------------------------
// Banded Cholesky direct solve of a symmetric positive definite CSR matrix (not in the
// reference; the independent pin for CG and the nonlinear solves).  Returns 0 on success,
// the failing pivot row + 1 if the matrix is not positive definite.
// ---------------------------------------------------------------------------------------
ORC_API int orc_band_cholesky_solve(int n, const int* offsets, const int* cols, const double* vals, const double* b,
                                    double* xo) {
    int bw = 0;
    for (int r = 0; r < n; r++)
        for (int k = offsets[r]; k < offsets[r + 1]; k++) bw = std::max(bw, std::abs(cols[k] - r));
    const int W = bw + 1;
    // lower band storage: Lb[r*W + (r - c)] for c in [r-bw, r]
    std::vector<double> Lb((size_t)n * W, 0.0);
    for (int r = 0; r < n; r++)
        for (int k = offsets[r]; k < offsets[r + 1]; k++) {
            int c = cols[k];
            if (c <= r) Lb[(size_t)r * W + (r - c)] = vals[k];
        }
    for (int j = 0; j < n; j++) {
        double s = Lb[(size_t)j * W];
        int c0 = std::max(0, j - bw);
        for (int k = c0; k < j; k++) {
            double l = Lb[(size_t)j * W + (j - k)];
            s -= l * l;
        }
        if (!(s > 0.0)) return j + 1;
        double ljj = std::sqrt(s);
        Lb[(size_t)j * W] = ljj;
        int r1 = std::min(n - 1, j + bw);
#pragma omp parallel for schedule(static) if (r1 - j > 256)
        for (int i = j + 1; i <= r1; i++) {
            double t = Lb[(size_t)i * W + (i - j)];
            int k0 = std::max(std::max(0, i - bw), c0);
            for (int k = k0; k < j; k++) t -= Lb[(size_t)i * W + (i - k)] * Lb[(size_t)j * W + (j - k)];
            Lb[(size_t)i * W + (i - j)] = t / ljj;
        }
    }
    std::vector<double> yv(n);
    for (int i = 0; i < n; i++) {
        double t = b[i];
        for (int k = std::max(0, i - bw); k < i; k++) t -= Lb[(size_t)i * W + (i - k)] * yv[k];
        yv[i] = t / Lb[(size_t)i * W];
    }
    for (int i = n - 1; i >= 0; i--) {
        double t = yv[i];
        int r1 = std::min(n - 1, i + bw);
        for (int k = i + 1; k <= r1; k++) t -= Lb[(size_t)k * W + (k - i)] * xo[k];
        xo[i] = t / Lb[(size_t)i * W];
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// Implicit-Euler velocity solvers (SyncedNesterov / SyncedAdamW / SyncedAdamWNocoop).
// State lives in the caller's arrays; one call = one Solve() = one time step.
// ---------------------------------------------------------------------------------------
struct OrcModel {
    int nn, nq, n_elem, n_nodes;
    const int* conn;
    const double* grad;
    const double* detJ;
    const double* qw;
    double lambda, mu, eta_d, lambda_d;
    const int* m_off;  // node-level mass CSR
    const int* m_col;
    const double* m_val;
    const double* f_ext;
    int n_fixed;  // constraints = 3 * n_fixed
    const int* fixed;
    const double* xref;  // constraint targets (SoA, n_nodes each)
    const double* yref;
    const double* zref;
};

struct OrcState {
    double* x;  // element positions (SoA), updated in place
    double* y;
    double* z;
    double* v_guess;  // 3N
    double* v_prev;   // 3N
    double* lam;      // 3 * n_fixed
};

namespace {

struct Work {
    std::vector<double> xp, yp, zp, f, c, g;
};

void positions_from(const OrcModel& M, const OrcState& S, const Work& W, const double* v, double dt) {
    for (int i = 0; i < M.n_nodes; i++) {
        S.x[i] = W.xp[i] + dt * v[3 * i + 0];
        S.y[i] = W.yp[i] + dt * v[3 * i + 1];
        S.z[i] = W.zp[i] + dt * v[3 * i + 2];
    }
}

void constraints(const OrcModel& M, const OrcState& S, std::vector<double>& c) {
    for (int k = 0; k < M.n_fixed; k++) {
        int n = M.fixed[k];
        c[3 * k + 0] = S.x[n] - M.xref[n];
        c[3 * k + 1] = S.y[n] - M.yref[n];
        c[3 * k + 2] = S.z[n] - M.zref[n];
    }
}

// g = M (v_guess - v_prev)/dt + f_int - f_ext + dt J^T (lam + rho c).
// `src/solvers/SyncedNesterov.cu:32-95`.
void gradient(const OrcModel& M, const OrcState& S, Work& W, double dt, double rho) {
    const int N3 = 3 * M.n_nodes;
    const double inv_dt = 1.0 / dt;
    for (int t = 0; t < N3; t++) {
        int ni = t / 3, di = t % 3;
        double res = 0.0;
        for (int k = M.m_off[ni]; k < M.m_off[ni + 1]; k++) {
            int tj = M.m_col[k] * 3 + di;
            res += M.m_val[k] * (S.v_guess[tj] - S.v_prev[tj]) * inv_dt;
        }
        res -= (-W.f[t]);
        res -= M.f_ext[t];
        W.g[t] = res;
    }
    // J^T rows: DOF 3*fixed[k]+d gets constraint 3k+d (value 1.0); repeated fixed nodes
    // contribute once per occurrence, as the CSR built at FEAT10Data.cu:420-473 does.
    for (int k = 0; k < M.n_fixed; k++)
        for (int d = 0; d < 3; d++) {
            int ci = 3 * k + d;
            W.g[3 * M.fixed[k] + d] += dt * 1.0 * (S.lam[ci] + rho * W.c[ci]);
        }
}

void eval_force(const OrcModel& M, const OrcState& S, Work& W, const double* v) {
    orc_internal_force(M.nn, M.nq, M.n_elem, M.n_nodes, M.conn, S.x, S.y, S.z, M.grad, M.detJ, M.qw, M.lambda, M.mu,
                       M.eta_d, M.lambda_d, v, W.f.data());
}

double nrm(const std::vector<double>& a) { return std::sqrt(dot((int)a.size(), a.data(), a.data())); }
double nrm(int n, const double* a) { return std::sqrt(dot(n, a, a)); }

}  // namespace

// One SyncedNesterov step.  `src/solvers/SyncedNesterov.cu:97-360`.
// Returns the total number of inner iterations executed.
ORC_API int orc_nesterov_step(const OrcModel* Mp, const OrcState* Sp, double alpha, double rho, double inner_tol,
                              double outer_tol, int max_outer, int max_inner, double dt) {
    const OrcModel& M = *Mp;
    const OrcState& S = *Sp;
    const int N = M.n_nodes, N3 = 3 * N, NC = 3 * M.n_fixed;
    Work W;
    W.xp.assign(S.x, S.x + N);
    W.yp.assign(S.y, S.y + N);
    W.zp.assign(S.z, S.z + N);
    W.f.assign(N3, 0.0);
    W.c.assign(NC, 0.0);
    W.g.assign(N3, 0.0);
    std::vector<double> vk(N3), vkm1(N3), vnext(N3), yv(N3);
    int total = 0;
    bool outer_flag = false;
    for (int outer = 0; outer < max_outer; outer++) {
        if (outer_flag) continue;
        double t = 1.0;
        for (int i = 0; i < N3; i++) vk[i] = vkm1[i] = S.v_guess[i];
        double prev_norm_g = 0.0;
        bool inner_flag = false;
        for (int inner = 0; inner < max_inner; inner++) {
            if (inner_flag) continue;
            total++;
            double t_next = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
            double beta = (t - 1.0) / t_next;
            for (int i = 0; i < N3; i++) {
                yv[i] = vk[i] + beta * (vk[i] - vkm1[i]);
                S.v_guess[i] = yv[i];
            }
            positions_from(M, S, W, S.v_guess, dt);
            eval_force(M, S, W, S.v_guess);
            constraints(M, S, W.c);
            gradient(M, S, W, dt, rho);
            double norm_g = nrm(W.g);
            if (inner > 0 && std::fabs(norm_g - prev_norm_g) < inner_tol) inner_flag = true;
            for (int i = 0; i < N3; i++) vnext[i] = yv[i] - alpha * W.g[i];
            double nvn = nrm(vnext), nvk = nrm(vk);
            if (inner > 0 && std::fabs(nvn - nvk) < inner_tol) inner_flag = true;
            for (int i = 0; i < N3; i++) {
                vkm1[i] = vk[i];
                vk[i] = vnext[i];
                S.v_guess[i] = vnext[i];
            }
            t = t_next;
            prev_norm_g = norm_g;
        }
        for (int i = 0; i < N3; i++) S.v_prev[i] = S.v_guess[i];
        positions_from(M, S, W, S.v_guess, dt);
        constraints(M, S, W.c);
        for (int i = 0; i < NC; i++) S.lam[i] += rho * dt * W.c[i];
        if (nrm(NC, W.c.data()) < outer_tol) outer_flag = true;
    }
    positions_from(M, S, W, S.v_guess, dt);
    return total;
}

// One SyncedAdamW step (cooperative variant).  `src/solvers/SyncedAdamW.cu:99-344`.
// With nocoop != 0 it follows SyncedAdamWNocoop instead (`SyncedAdamWNocoop.cu:255-436`):
// lr_k = lr0 * decay^(k+1) and bias factors 1/(1-beta^t) from the host, max_outer forced
// to 1 without constraints, the outer stop also requires an inner convergence.
ORC_API int orc_adamw_step(const OrcModel* Mp, const OrcState* Sp, double lr0, double beta1, double beta2, double eps,
                           double weight_decay, double lr_decay, double inner_tol, double outer_tol, double rho,
                           int max_outer, int max_inner, double dt, int check_interval, double inner_rtol,
                           int nocoop) {
    const OrcModel& M = *Mp;
    const OrcState& S = *Sp;
    const int N = M.n_nodes, N3 = 3 * N, NC = 3 * M.n_fixed;
    Work W;
    W.xp.assign(S.x, S.x + N);
    W.yp.assign(S.y, S.y + N);
    W.zp.assign(S.z, S.z + N);
    W.f.assign(N3, 0.0);
    W.c.assign(NC, 0.0);
    W.g.assign(N3, 0.0);
    std::vector<double> m(N3), va(N3);
    int total = 0;
    if (nocoop && check_interval <= 0) check_interval = 1;
    int eff_outer = (nocoop && NC == 0) ? 1 : max_outer;
    bool outer_flag = false;
    for (int outer = 0; outer < eff_outer; outer++) {
        if (outer_flag) {
            if (nocoop) break;
            continue;
        }
        double lr = lr0, t = 1.0;
        std::fill(m.begin(), m.end(), 0.0);
        std::fill(va.begin(), va.end(), 0.0);
        std::fill(W.g.begin(), W.g.end(), 0.0);
        double norm_g0 = -1.0;
        bool inner_flag = false;
        for (int inner = 0; inner < max_inner; inner++) {
            if (inner_flag) {
                if (nocoop) break;
                continue;
            }
            total++;
            double lr_k, inv1, inv2;
            if (nocoop) {
                lr_k = lr0 * std::pow(lr_decay, inner + 1);
                double tc = (double)(inner + 2);
                inv1 = 1.0 / (1.0 - std::pow(beta1, tc));
                inv2 = 1.0 / (1.0 - std::pow(beta2, tc));
            } else {
                lr = lr * lr_decay;
                t += 1;
                lr_k = lr;
                inv1 = inv2 = 0.0;
            }
            for (int i = 0; i < N3; i++) {
                double gi = W.g[i], vg = S.v_guess[i];
                if (nocoop) {
                    m[i] = beta1 * m[i] + (1.0 - beta1) * gi;
                    va[i] = beta2 * va[i] + (1.0 - beta2) * gi * gi;
                    double mh = m[i] * inv1, vh = va[i] * inv2;
                    S.v_guess[i] = vg - lr_k * (mh / (std::sqrt(vh) + eps) + weight_decay * vg);
                } else {
                    m[i] = beta1 * m[i] + (1 - beta1) * gi;
                    va[i] = beta2 * va[i] + (1 - beta2) * gi * gi;
                    double mh = m[i] / (1 - std::pow(beta1, t));
                    double vh = va[i] / (1 - std::pow(beta2, t));
                    S.v_guess[i] = vg - lr_k * (mh / (std::sqrt(vh) + eps) + weight_decay * vg);
                }
            }
            positions_from(M, S, W, S.v_guess, dt);
            eval_force(M, S, W, S.v_guess);
            if (NC > 0) constraints(M, S, W.c);
            gradient(M, S, W, dt, rho);
            if (inner % check_interval == 0) {
                double ng = nrm(W.g);
                if (norm_g0 < 0.0) norm_g0 = ng;
                double nv = nrm(N3, S.v_guess);
                double tol_abs = inner_tol * (1.0 + nv);
                double tol_rel = (inner_rtol > 0.0 && norm_g0 > 0.0) ? inner_rtol * norm_g0 : 0.0;
                if (ng <= tol_abs || (tol_rel > 0.0 && ng <= tol_rel)) inner_flag = true;
            }
        }
        for (int i = 0; i < N3; i++) S.v_prev[i] = S.v_guess[i];
        positions_from(M, S, W, S.v_guess, dt);
        if (nocoop) {
            if (NC > 0) {
                constraints(M, S, W.c);
                for (int i = 0; i < NC; i++) S.lam[i] += rho * dt * W.c[i];
                if (nrm(NC, W.c.data()) < outer_tol && inner_flag) outer_flag = true;
            }
        } else {
            constraints(M, S, W.c);
            for (int i = 0; i < NC; i++) S.lam[i] += rho * dt * W.c[i];
            if (nrm(NC, W.c.data()) < outer_tol) outer_flag = true;
        }
    }
    positions_from(M, S, W, S.v_guess, dt);
    return total;
}

// Step gradient g(v) at the given state (positions are x_prev + dt v), used by tests to
// state residuals of the implicit-Euler equation.  xp/yp/zp = positions at step start.
ORC_API void orc_step_gradient(const OrcModel* Mp, const double* xp, const double* yp, const double* zp,
                               const double* v, const double* v_prev, const double* lam, double dt, double rho,
                               double* g_out) {
    const OrcModel& M = *Mp;
    const int N = M.n_nodes, N3 = 3 * N, NC = 3 * M.n_fixed;
    std::vector<double> x(N), y(N), z(N), vg(v, v + N3), vp(v_prev, v_prev + N3), lm(lam, lam + NC);
    OrcState S{x.data(), y.data(), z.data(), vg.data(), vp.data(), lm.data()};
    Work W;
    W.xp.assign(xp, xp + N);
    W.yp.assign(yp, yp + N);
    W.zp.assign(zp, zp + N);
    W.f.assign(N3, 0.0);
    W.c.assign(NC, 0.0);
    W.g.assign(N3, 0.0);
    positions_from(M, S, W, vg.data(), dt);
    eval_force(M, S, W, vg.data());
    constraints(M, S, W.c);
    gradient(M, S, W, dt, rho);
    std::copy(W.g.begin(), W.g.end(), g_out);
}

ORC_API int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

ORC_API void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// ---------------------------------------------------------------------------------------
// Material selection and the Mooney-Rivlin point functions (F, P column-major 3x3,
// A[i + 3 j + 9 k + 27 l] = dP_ij/dF_kl).
// ---------------------------------------------------------------------------------------
ORC_API void orc_set_material(int model, double mu10, double mu01, double kappa) {
    g_model = model;
    g_mu10 = mu10;
    g_mu01 = mu01;
    g_kappa = kappa;
}

ORC_API void orc_mr_point(const double* F9, double mu10, double mu01, double kappa, double* W, double* P9,
                          double* A81) {
    double F[3][3], P[3][3], A[3][3][3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) F[i][j] = F9[i + 3 * j];
    if (W) *W = mr_W(F, mu10, mu01, kappa);
    if (P9) {
        mr_P(F, mu10, mu01, kappa, P);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) P9[i + 3 * j] = P[i][j];
    }
    if (A81) {
        mr_A(F, mu10, mu01, kappa, A);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) A81[i + 3 * j + 9 * k + 27 * l] = A[i][j][k][l];
    }
}
