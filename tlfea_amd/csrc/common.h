// Shared helpers of the gfx950 TLFEA library: error plumbing, launch helpers, small
// fixed-size tensor algebra in registers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/tlfea_hip.h"

namespace tlfea_hip {

constexpr int TILE = TLFEA_TILE;

// Thread-local message of the last failing call, returned by tlfea_last_error().
void set_error(const char* fmt, ...);

#define TLFEA_HIP_TRY(expr)                                                                       \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            ::tlfea_hip::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                                   __LINE__);                                                     \
            return TLFEA_EHIP;                                                                    \
        }                                                                                         \
    } while (0)

#define TLFEA_REQUIRE(cond, ...)                 \
    do {                                         \
        if (!(cond)) {                           \
            ::tlfea_hip::set_error(__VA_ARGS__); \
            return TLFEA_EINVAL;                 \
        }                                        \
    } while (0)

// Checks the launch that was just issued.
#define TLFEA_LAUNCH_CHECK() TLFEA_HIP_TRY(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t padded(int64_t n) { return (n + TILE - 1) / TILE * TILE; }

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// The dispatcher hands workgroup b to XCD b % 8, and each of the 8 XCDs has its own L2.
// xcd_block(b, n) is a bijection of [0, n) that gives XCD x one contiguous run of logical
// blocks (the first n % 8 XCDs one block more), so the node data a run gathers stays in one
// L2 instead of being fetched by all eight.  Results never depend on it: every logical
// block is still run exactly once.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned n) {
    constexpr unsigned X = 8;
    const unsigned q = n / X, r = n % X, x = b % X, k = b / X;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// ------------------------------------------------------------------------------------
// Device algebra
// ------------------------------------------------------------------------------------
// P = lambda (tr(F^T F)/2 - 3/2) F + mu (F F^T F - F)  (SVK, src/materials/SVK.cuh:17-39)
template <typename T>
__device__ __forceinline__ void svk_stress(const T F[3][3], T lambda, T mu, T P[3][3]) {
    T FFt[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) FFt[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
    // tr(F^T F) == tr(F F^T)
    const T tr = FFt[0][0] + FFt[1][1] + FFt[2][2];
    const T lf = lambda * (T(0.5) * tr - T(1.5));
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const T fftf = FFt[i][0] * F[0][j] + FFt[i][1] * F[1][j] + FFt[i][2] * F[2][j];
            P[i][j] = lf * F[i][j] + mu * (fftf - F[i][j]);
        }
}

// Kelvin-Voigt: Edot = (Fdot^T F + F^T Fdot)/2, S = 2 eta Edot + ld tr(Edot) I, P_vis = F S
// (src/elements/FEAT10DataFunc.cuh:135-218)
template <typename T>
__device__ __forceinline__ void kv_stress(const T F[3][3], const T Fd[3][3], T eta, T ld, T Pv[3][3]) {
    T S[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            T a = Fd[0][i] * F[0][j] + Fd[1][i] * F[1][j] + Fd[2][i] * F[2][j];
            T b = F[0][i] * Fd[0][j] + F[1][i] * Fd[1][j] + F[2][i] * Fd[2][j];
            S[i][j] = T(0.5) * (a + b);
        }
    const T tr = S[0][0] + S[1][1] + S[2][2];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] = T(2) * eta * S[i][j] + (i == j ? ld * tr : T(0));
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Pv[i][j] = F[i][0] * S[0][j] + F[i][1] * S[1][j] + F[i][2] * S[2][j];
}

// Compressible Mooney-Rivlin (src/materials/MooneyRivlin.cuh:45-227): the derivative of
//     W = mu10 (J^-2/3 I1 - 3) + mu01 (J^-4/3 I2 - 3) + kappa/2 (J - 1)^2,
//     P = t1 (F - I1/3 F^-T) + t2 (I1 F - F C - 2 I2/3 F^-T) + t3 F^-T,
// t1 = 2 mu10 J^-2/3, t2 = 2 mu01 J^-4/3, t3 = kappa (J - 1) J, C = F^T F,
// I2 = (I1^2 - tr C^2)/2.  F^-T uses det F clamped to +-1e-12 in magnitude, as the
// reference does.
struct MrKin {
    double C[3][3], FC[3][3], FinvT[3][3], I1, I2, J, t1, t2, t3;
};

__device__ __forceinline__ void mr_kinematics(const double F[3][3], double mu10, double mu01, double kappa,
                                              MrKin& k) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) k.C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    k.I1 = k.C[0][0] + k.C[1][1] + k.C[2][2];
    double trC2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) trC2 += k.C[i][j] * k.C[j][i];
    k.I2 = 0.5 * (k.I1 * k.I1 - trC2);
    k.J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
          F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    double d = k.J;
    if (fabs(d) < 1e-12) d = d >= 0.0 ? 1e-12 : -1e-12;
    const double id = 1.0 / d;
    // F^-T = cofactor(F) / det F
    k.FinvT[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * id;
    k.FinvT[0][1] = (F[1][2] * F[2][0] - F[1][0] * F[2][2]) * id;
    k.FinvT[0][2] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * id;
    k.FinvT[1][0] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * id;
    k.FinvT[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * id;
    k.FinvT[1][2] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * id;
    k.FinvT[2][0] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * id;
    k.FinvT[2][1] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * id;
    k.FinvT[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * id;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) k.FC[i][j] = F[i][0] * k.C[0][j] + F[i][1] * k.C[1][j] + F[i][2] * k.C[2][j];
    const double j13 = cbrt(k.J);
    const double jm23 = 1.0 / (j13 * j13);
    k.t1 = 2.0 * mu10 * jm23;
    k.t2 = 2.0 * mu01 * jm23 * jm23;
    k.t3 = kappa * (k.J - 1.0) * k.J;
}

__device__ __forceinline__ void mr_stress(const double F[3][3], double mu10, double mu01, double kappa,
                                          double P[3][3]) {
    MrKin k;
    mr_kinematics(F, mu10, mu01, kappa, k);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            P[i][j] = k.t1 * (F[i][j] - k.I1 / 3.0 * k.FinvT[i][j]) +
                      k.t2 * (k.I1 * F[i][j] - k.FC[i][j] - 2.0 * k.I2 / 3.0 * k.FinvT[i][j]) + k.t3 * k.FinvT[i][j];
}

// A[i][j][k][l] = dP_ij / dF_kl of mr_stress (MooneyRivlin.cuh:106-227), written as the
// chain rule of the three terms: d(t1) = -2/3 t1 F^-T_kl, d(t2) = -4/3 t2 F^-T_kl,
// d(t3) = kappa (2J - 1) J F^-T_kl, d(F^-T_ij) = -F^-T_il F^-T_kj, d(I1) = 2 F_kl,
// d(I2) = 2 (I1 F - F C)_kl, d(FC)_ij = delta_ik C_lj + F_il F_kj + delta_jl (F F^T)_ik.
__device__ __forceinline__ double mr_tangent_entry(const double F[3][3], const double FFt[3][3], const MrKin& k,
                                                   double kappa, int i, int j, int kk, int l) {
    const double dik = i == kk ? 1.0 : 0.0, djl = j == l ? 1.0 : 0.0;
    const double fk = k.FinvT[kk][l];
    const double term1 = F[i][j] - k.I1 / 3.0 * k.FinvT[i][j];
    const double term2 = k.I1 * F[i][j] - k.FC[i][j] - 2.0 * k.I2 / 3.0 * k.FinvT[i][j];
    const double dinv = -k.FinvT[i][l] * k.FinvT[kk][j];
    const double dT1 = dik * djl - 2.0 / 3.0 * F[kk][l] * k.FinvT[i][j] - k.I1 / 3.0 * dinv;
    const double dT2 = 2.0 * F[kk][l] * F[i][j] + k.I1 * dik * djl -
                       (dik * k.C[l][j] + F[i][l] * F[kk][j] + djl * FFt[i][kk]) -
                       4.0 / 3.0 * (k.I1 * F[kk][l] - k.FC[kk][l]) * k.FinvT[i][j] - 2.0 * k.I2 / 3.0 * dinv;
    return -2.0 / 3.0 * k.t1 * fk * term1 + k.t1 * dT1 - 4.0 / 3.0 * k.t2 * fk * term2 + k.t2 * dT2 +
           kappa * (2.0 * k.J - 1.0) * k.J * fk * k.FinvT[i][j] + k.t3 * dinv;
}

// dP = A : dF of mr_stress at F for a direction dF, without forming A: the derivative of
// P = t1 T1 + t2 T2 + t3 G (T1 = F - I1/3 G, T2 = I1 F - F C - 2 I2/3 G, G = F^-T) with
// dJ = J G:dF, dG = -G dF^T G, dI1 = 2 F:dF, dI2 = 2 (I1 F - F C):dF,
// d(F C) = dF C + F (dF^T F + F^T dF), dt1 = -2/3 t1 G:dF, dt2 = -4/3 t2 G:dF,
// dt3 = kappa (2J - 1) J G:dF -- the contraction of mr_tangent_entry with dF.
__device__ __forceinline__ void mr_directional(const double F[3][3], const double dF[3][3], double mu10,
                                               double mu01, double kappa, double dP[3][3]) {
    MrKin k;
    mr_kinematics(F, mu10, mu01, kappa, k);
    double gd = 0.0, dI1 = 0.0, dI2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            gd += k.FinvT[i][j] * dF[i][j];
            dI1 += 2.0 * F[i][j] * dF[i][j];
            dI2 += 2.0 * (k.I1 * F[i][j] - k.FC[i][j]) * dF[i][j];
        }
    double W[3][3], dC[3][3];  // W = dF^T G, dC = dF^T F + F^T dF
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double w = 0.0, c = 0.0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                w += dF[a][i] * k.FinvT[a][j];
                c += dF[a][i] * F[a][j] + F[a][i] * dF[a][j];
            }
            W[i][j] = w;
            dC[i][j] = c;
        }
    const double dt1 = -2.0 / 3.0 * k.t1 * gd, dt2 = -4.0 / 3.0 * k.t2 * gd, dt3 = kappa * (2.0 * k.J - 1.0) * k.J * gd;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double dG = 0.0, dFC = 0.0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                dG -= k.FinvT[i][a] * W[a][j];
                dFC += dF[i][a] * k.C[a][j] + F[i][a] * dC[a][j];
            }
            const double G = k.FinvT[i][j];
            const double T1 = F[i][j] - k.I1 / 3.0 * G;
            const double T2 = k.I1 * F[i][j] - k.FC[i][j] - 2.0 * k.I2 / 3.0 * G;
            dP[i][j] = dt1 * T1 + k.t1 * (dF[i][j] - dI1 / 3.0 * G - k.I1 / 3.0 * dG) + dt2 * T2 +
                       k.t2 * (dI1 * F[i][j] + k.I1 * dF[i][j] - dFC - 2.0 * dI2 / 3.0 * G - 2.0 * k.I2 / 3.0 * dG) +
                       dt3 * G + k.t3 * dG;
        }
}

// P of the element's material: SVK (model 0) or Mooney-Rivlin (model 1).  The hot force
// kernels take the model as a template parameter (a runtime branch costs the SVK variant
// its register allocation); k_calc_p decides at run time.
template <bool MR>
__device__ __forceinline__ void material_stress_t(const double F[3][3], const tlfea_material& m, double P[3][3]) {
    if constexpr (MR)
        mr_stress(F, m.mu10, m.mu01, m.kappa, P);
    else
        svk_stress(F, m.lambda, m.mu, P);
}
__device__ __forceinline__ void material_stress(const double F[3][3], const tlfea_material& m, double P[3][3]) {
    if (m.model == 1)
        mr_stress(F, m.mu10, m.mu01, m.kappa, P);
    else
        svk_stress(F, m.lambda, m.mu, P);
}

// Material models this library evaluates (tlfea_material.model).
#define TLFEA_REQUIRE_MODEL(mat) \
    TLFEA_REQUIRE((mat).model == 0 || (mat).model == 1, "unknown material model %d (0 SVK, 1 Mooney-Rivlin)", (mat).model)

__device__ __forceinline__ double det3(const double J[3][3]) {
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

}  // namespace tlfea_hip
