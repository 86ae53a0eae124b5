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

__device__ __forceinline__ double det3(const double J[3][3]) {
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

}  // namespace tlfea_hip
