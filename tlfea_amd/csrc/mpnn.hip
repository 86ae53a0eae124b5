// Learned warm start (not in the reference): a message-passing network on the mesh node
// graph predicts the linear-static displacement of a cantilever from its rest geometry,
// boundary conditions, load and material, as the initial guess of the CG solve.
//
// Network (BF16 weights and activations, FP32 accumulation):
//   h0 = bf16(relu(x W_in^T + b_in))                              x: [n, 16] features
//   for l = 1..L:  m = bf16(mean_{j in N(i)} h_j)                  (N(i) includes i)
//                  pre_l = [h, m] W_l^T + b_l                     (FP32)
//                  h = bf16(h + relu(pre_l))
//   out = h W_out^T + b_out                                       (FP32, 3 used of 16)
// The GEMMs run on the matrix cores (v_mfma_f32_16x16x32_bf16, gfx950: 16x16 tiles, K = 32
// per instruction, 8 BF16 operands and 4 FP32 results per lane; operand fragments are one
// 16-byte load per lane).  Backward: the same layout, with dpre rounded to
// BF16 before each GEMM; weight gradients are reduced over nodes in per-block FP32
// partials summed in a fixed order.  Optimiser: Adam on FP32 master weights.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace tlfea_hip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int MW = 64;     // hidden width
constexpr int MFEAT = 16;  // input features (padded)
constexpr int MOUT = 16;   // output rows (3 used, padded to one MFMA tile)
constexpr int WG_ROWS = 128;    // nodes per weight-gradient chunk (staged in LDS)
constexpr int WG_BLOCKS = 1024; // at most this many weight-gradient blocks (partials)

// round to nearest even (finite inputs)
__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = __builtin_bit_cast(__bf16, (unsigned short)0);
    return v;
}
// 8 consecutive BF16 values (16-byte aligned)
__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// Y[M, N] = sum_p A_p[M, K_p] B_p, with B_p given transposed as BT_p[n * ldb + k] (row n of
// B^T is column n of B).  One wave computes 16 rows x N columns; 4 waves per block.  Per
// 32-wide k step lane (li, kg) holds k = 8 kg .. 8 kg + 7 of A's row li and of B^T's row li of
// each 16-column tile (K a multiple of 16: the k >= K half of the last step is zero; rows and
// ldb multiples of 8, so every fragment is one aligned 16-byte load).
// Epilogue (mode):
//   0: y[r * ldy + c] = acc + (bias ? bias[c] : 0)                      (FP32)
//   1: pre = acc + bias -> pre[r * N + c] (FP32);
//      h_out[r * N + c] = bf16((h_in ? h_in[r * N + c] : 0) + relu(pre))
template <int N>
__global__ __launch_bounds__(256) void k_gemm(int M, const bf16* __restrict__ A0, int K0,
                                              const bf16* __restrict__ BT0, int ldb0, const bf16* __restrict__ A1,
                                              int K1, const bf16* __restrict__ BT1, int ldb1,
                                              const float* __restrict__ bias, int mode, const bf16* __restrict__ h_in,
                                              float* __restrict__ pre, bf16* __restrict__ h_out,
                                              float* __restrict__ y, int ldy) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + wave) * 16;
    if (r0 >= M) return;
    const int li = lane & 15, kg = lane >> 4;
    f32x4 acc[N / 16];
#pragma unroll
    for (int t = 0; t < N / 16; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int part = 0; part < 2; part++) {
        const bf16* A = part ? A1 : A0;
        const bf16* BT = part ? BT1 : BT0;
        const int K = part ? K1 : K0;
        const int ldb = part ? ldb1 : ldb0;
        if (!A) continue;
        const int row = r0 + li;
        for (int k0 = 0; k0 < K; k0 += 32) {
            const int kk = k0 + 8 * kg;
            const bool kin = kk < K;
            const bf16x8 a = row < M && kin ? load8(A + (int64_t)row * K + kk) : zero8();
#pragma unroll
            for (int t = 0; t < N / 16; t++) {
                const bf16x8 bv = kin ? load8(BT + (int64_t)(t * 16 + li) * ldb + kk) : zero8();
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
            }
        }
    }
    // acc[t][j] = Y[r0 + 4*kg + j][t*16 + li]
#pragma unroll
    for (int t = 0; t < N / 16; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = r0 + 4 * kg + j, c = t * 16 + li;
            if (r >= M) continue;
            const float v = acc[t][j] + (bias ? bias[c] : 0.f);
            if (mode == 0) {
                y[(int64_t)r * ldy + c] = v;
            } else {
                pre[(int64_t)r * N + c] = v;
                const float base = h_in ? bf2f(h_in[(int64_t)r * N + c]) : 0.f;
                h_out[(int64_t)r * N + c] = f2bf(base + fmaxf(v, 0.f));
            }
        }
}

// Weight gradient partials: P[blk][n * K + k] = sum_{m in the block's chunks} dY[m, n] A[m, k]
// (dY: [M, N] bf16, A: [M, K] bf16).  Block b takes the WG_ROWS-node chunks b, b + grid,
// b + 2 grid, ... in order, each staged in LDS, and keeps its (at most four per wave) output
// tiles in registers across them: few partials, in a fixed summation order.
template <int N, int K>
__global__ __launch_bounds__(256) void k_wgrad(int M, const bf16* __restrict__ dY, const bf16* __restrict__ A,
                                               float* __restrict__ P) {
    __shared__ bf16 sY[WG_ROWS * N];
    __shared__ bf16 sA[WG_ROWS * K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, kg = lane >> 4;
    constexpr int TN = N / 16, TK = K / 16, NT = (TN * TK + 3) / 4;
    f32x4 acc[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nchunks = (M + WG_ROWS - 1) / WG_ROWS;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int m0 = c * WG_ROWS;
        __syncthreads();  // the previous chunk's fragments have been read
        for (int t = threadIdx.x; t < WG_ROWS * N; t += 256) {
            const int m = m0 + t / N;
            sY[t] = m < M ? dY[(int64_t)m0 * N + t] : (bf16)0;
        }
        for (int t = threadIdx.x; t < WG_ROWS * K; t += 256) {
            const int m = m0 + t / K;
            sA[t] = m < M ? A[(int64_t)m0 * K + t] : (bf16)0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NT; q++) {
            const int tile = wave + 4 * q;
            if (tile >= TN * TK) break;
            const int tn = tile / TK, tk = tile % TK;
            for (int mm = 0; mm < WG_ROWS; mm += 32) {
                // MFMA A operand: rows n, reduction m -> dY^T[tn*16 + li][mm + 8kg + j]
                // MFMA B operand: reduction m, cols k -> A[mm + 8kg + j][tk*16 + li]
                bf16x8 a, b;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    a[j] = __builtin_bit_cast(__bf16, sY[(mm + 8 * kg + j) * N + tn * 16 + li]);
                    b[j] = __builtin_bit_cast(__bf16, sA[(mm + 8 * kg + j) * K + tk * 16 + li]);
                }
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int tile = wave + 4 * q;
        if (tile >= TN * TK) break;
        const int tn = tile / TK, tk = tile % TK;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = tn * 16 + 4 * kg + j, k = tk * 16 + li;
            P[(int64_t)blockIdx.x * N * K + n * K + k] = acc[q][j];
        }
    }
}

// g[(i / cols) * ldg + i % cols] += sum_b P[b][i]  (fixed order; len = rows * cols)
__global__ void k_sum_partials(int nb, int len, int cols, int ldg, const float* __restrict__ P,
                               float* __restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.f;
#pragma unroll 8
    for (int b = 0; b < nb; b++) s += P[(int64_t)b * len + i];
    g[(i / cols) * ldg + i % cols] += s;
}

// column sums of an FP32 [M, N] matrix -> P[blk][N], block b summing the WG_ROWS-row chunks
// b, b + grid, ... in order
__global__ void k_colsum_partials(int M, int N, const float* __restrict__ Y, float* __restrict__ P) {
    const int c = threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (int m0 = blockIdx.x * WG_ROWS; m0 < M; m0 += gridDim.x * WG_ROWS)
        for (int m = m0; m < min(M, m0 + WG_ROWS); m++) s += Y[(int64_t)m * N + c];
    P[(int64_t)blockIdx.x * N + c] = s;
}

// m_i = bf16(mean_{j in row i} h_j) over the node graph (CSR includes the diagonal).  32
// lanes per node (two nodes per wave), two features per lane (one 4-byte load per
// neighbour row), four neighbour rows in flight; neighbours are added in CSR order.
__global__ __launch_bounds__(256) void k_agg(int n, const int* __restrict__ off, const int* __restrict__ col,
                                             const bf16* __restrict__ h, bf16* __restrict__ m) {
    const int node = blockIdx.x * 8 + (threadIdx.x >> 5), f2 = threadIdx.x & 31;
    if (node >= n) return;
    const int a = off[node], b = off[node + 1];
    const uint32_t* h2 = reinterpret_cast<const uint32_t*>(h);
    float s0 = 0.f, s1 = 0.f;
    int k = a;
    for (; k + 4 <= b; k += 4) {
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = h2[(int64_t)col[k + u] * (MW / 2) + f2];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            s0 += bf2f((bf16)(v[u] & 0xffffu));
            s1 += bf2f((bf16)(v[u] >> 16));
        }
    }
    for (; k < b; k++) {
        const uint32_t v = h2[(int64_t)col[k] * (MW / 2) + f2];
        s0 += bf2f((bf16)(v & 0xffffu));
        s1 += bf2f((bf16)(v >> 16));
    }
    const float deg = (float)(b - a);
    reinterpret_cast<uint32_t*>(m)[(int64_t)node * (MW / 2) + f2] =
        (uint32_t)f2bf(s0 / deg) | ((uint32_t)f2bf(s1 / deg) << 16);
}

// deg(i) per node (as float)
__global__ void k_deg(int n, const int* __restrict__ off, float* __restrict__ deg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) deg[i] = (float)(off[i + 1] - off[i]);
}

// dh_j += sum_{i in row j} dm_i / deg(i)  (the graph is symmetric); 32 lanes per node, two
// features per lane, four neighbour rows in flight, CSR order
__global__ __launch_bounds__(256) void k_agg_t(int n, const int* __restrict__ off, const int* __restrict__ col,
                                               const float* __restrict__ deg, const float* __restrict__ dm,
                                               float* __restrict__ dh) {
    const int node = blockIdx.x * 8 + (threadIdx.x >> 5), f2 = threadIdx.x & 31;
    if (node >= n) return;
    const float2* dm2 = reinterpret_cast<const float2*>(dm);
    float s0 = 0.f, s1 = 0.f;
    int k = off[node];
    const int e = off[node + 1];
    for (; k + 4 <= e; k += 4) {
        float2 v[4];
        float w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = col[k + u];
            v[u] = dm2[(int64_t)i * (MW / 2) + f2];
            w[u] = deg[i];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            s0 += v[u].x / w[u];
            s1 += v[u].y / w[u];
        }
    }
    for (; k < e; k++) {
        const int i = col[k];
        const float2 v = dm2[(int64_t)i * (MW / 2) + f2];
        s0 += v.x / deg[i];
        s1 += v.y / deg[i];
    }
    float2* o = reinterpret_cast<float2*>(dh) + (int64_t)node * (MW / 2) + f2;
    float2 cur = *o;
    cur.x += s0;
    cur.y += s1;
    *o = cur;
}

// dpre = dh * (pre > 0) in FP32 and BF16
__global__ void k_relu_bwd(int64_t len, const float* __restrict__ dh, const float* __restrict__ pre,
                           float* __restrict__ dpre, bf16* __restrict__ dpre16) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const float v = pre[i] > 0.f ? dh[i] : 0.f;
    dpre[i] = v;
    dpre16[i] = f2bf(v);
}

__global__ void k_add_f32(int64_t len, const float* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) b[i] += a[i];
}

// loss = sum_i w_i |out_i - t_i|^2 / wsum over the 3 used outputs; dout (bf16, [n, 16],
// padded columns 0) = 2 w_i (out - t) / wsum; per-block loss partials
__global__ __launch_bounds__(256) void k_loss(int n, const float* __restrict__ out, const float* __restrict__ t,
                                              const float* __restrict__ w, float inv_wsum, bf16* __restrict__ dout,
                                              float* __restrict__ dout32, double* __restrict__ lpart) {
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        const float wi = w[i];
#pragma unroll
        for (int c = 0; c < MOUT; c++) {
            float g = 0.f;
            if (c < 3) {
                const float d = out[(int64_t)i * MOUT + c] - t[(int64_t)i * 3 + c];
                l += (double)wi * d * d;
                g = 2.f * wi * d * inv_wsum;
            }
            dout[(int64_t)i * MOUT + c] = f2bf(g);
            dout32[(int64_t)i * MOUT + c] = g;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0] * inv_wsum;
}

__global__ void k_adam(int len, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, const unsigned char* __restrict__ frozen, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len || (frozen && frozen[i])) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// bf16 copies of the weights: straight (forward) and transposed (input-gradient GEMMs)
__global__ void k_to_bf16(int len, const float* __restrict__ a, bf16* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) b[i] = f2bf(a[i]);
}
__global__ void k_to_bf16_t(int rows, int cols, const float* __restrict__ a, bf16* __restrict__ bt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    bt[c * rows + r] = f2bf(a[i]);
}

// Node features and per-instance displacement scales for the beams of a batch:
//   [ (X - lo)/ext (3), fixed, f/|F| (3), F/|F| (3), (log10 E - 8)/2, (nu - 0.325)/0.125,
//     log(ext_x/ext_y), log(ext_x/ext_z), xi^2, (3 xi^2 - xi^3)/2 ]   (xi = (X - lo_x)/ext_x,
//   the last two the cantilever deflection shapes of a clamped x = lo_x end)
// inst_par[i*12]: lo(3), ext(3), F(3), E, nu, unused.
__global__ void k_features(int n, const int* __restrict__ node_inst, const double* __restrict__ X,
                           const double* __restrict__ Y, const double* __restrict__ Z,
                           const unsigned char* __restrict__ dof_mask, const double* __restrict__ f,
                           const double* __restrict__ inst_par, bf16* __restrict__ feat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* P = inst_par + (int64_t)node_inst[i] * 12;
    const double Fn = sqrt(P[6] * P[6] + P[7] * P[7] + P[8] * P[8]);
    const double iF = Fn > 0.0 ? 1.0 / Fn : 0.0;
    float v[MFEAT];
    v[0] = (float)((X[i] - P[0]) / P[3]);
    v[1] = (float)((Y[i] - P[1]) / P[4]);
    v[2] = (float)((Z[i] - P[2]) / P[5]);
    v[3] = dof_mask[3 * (int64_t)i] ? 1.f : 0.f;
    for (int d = 0; d < 3; d++) v[4 + d] = (float)(f[3 * (int64_t)i + d] * iF);
    for (int d = 0; d < 3; d++) v[7 + d] = (float)(P[6 + d] * iF);
    v[10] = (float)((log10(P[9]) - 8.0) / 2.0);
    v[11] = (float)((P[10] - 0.325) / 0.125);
    v[12] = (float)log(P[3] / P[4]);
    v[13] = (float)log(P[3] / P[5]);
    const double xi = (X[i] - P[0]) / P[3];
    v[14] = (float)(xi * xi);
    v[15] = (float)(0.5 * (3.0 * xi * xi - xi * xi * xi));
    for (int c = 0; c < MFEAT; c++) feat[(int64_t)i * MFEAT + c] = f2bf(v[c]);
}

// dout (FP32 and BF16, [n, 16], padded columns 0) from a caller's FP32 [n, 3] gradient
__global__ void k_set_dout(int n, const float* __restrict__ g3, float* __restrict__ dout32, bf16* __restrict__ dout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < MOUT; c++) {
        const float g = c < 3 ? g3[(int64_t)i * 3 + c] : 0.f;
        dout32[(int64_t)i * MOUT + c] = g;
        dout[(int64_t)i * MOUT + c] = f2bf(g);
    }
}

// gradient of sum_inst w_inst e^T K e / E_inst w.r.t. the network outputs, where
// e = scale * out - u*: dout[i, d] = 2 w_inst scale[inst, d] (K e)[3i+d] / E_inst
__global__ void k_energy_dout(int n, const int* __restrict__ node_inst, const double* __restrict__ scale,
                              const double* __restrict__ Ke, const double* __restrict__ w_over_E,
                              float* __restrict__ g3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int inst = node_inst[i];
    for (int d = 0; d < 3; d++)
        g3[(int64_t)i * 3 + d] = (float)(2.0 * w_over_E[inst] * scale[3 * (int64_t)inst + d] * Ke[3 * (int64_t)i + d]);
}

// u0[3i+d] = scale[3 inst + d] * out[i, d] on free DOFs, 0 on fixed ones
__global__ void k_scale_out(int n, const int* __restrict__ node_inst, const float* __restrict__ out,
                            const double* __restrict__ scale, const unsigned char* __restrict__ dof_mask,
                            double* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* s = scale + 3 * (int64_t)node_inst[i];
    for (int d = 0; d < 3; d++) {
        const int64_t k = 3 * (int64_t)i + d;
        u[k] = dof_mask[k] ? 0.0 : s[d] * (double)out[(int64_t)i * MOUT + d];
    }
}

}  // namespace tlfea_hip

using namespace tlfea_hip;

// Parameter block (FP32, contiguous): W_in [64x16], b_in [64], then per layer W_l [64x128],
// b_l [64], then W_out [16x64] (rows 3..15 frozen at 0), b_out [16] (3..15 frozen).
struct tlfea_mpnn {
    int layers = 4;
    int n_params = 0;
    std::vector<int> off_W, off_b;  // [layers + 2]: encoder, layers, decoder
    float *d_w = nullptr, *d_g = nullptr, *d_m1 = nullptr, *d_m2 = nullptr;
    unsigned char* d_frozen = nullptr;
    bf16* d_w16 = nullptr;   // bf16 copy, same layout
    bf16* d_wt16 = nullptr;  // transposed bf16 copies of the W matrices (same offsets)
    int adam_t = 0;
    // workspace (sized for cap nodes)
    int cap = 0;
    bf16 *d_x = nullptr, *d_h = nullptr, *d_mg = nullptr, *d_dp16 = nullptr, *d_dout16 = nullptr;
    float *d_pre = nullptr, *d_out = nullptr, *d_dh = nullptr, *d_dm = nullptr, *d_dp = nullptr, *d_dout = nullptr;
    float* d_part = nullptr;
    double* d_lpart = nullptr;
    float* d_deg = nullptr;  // node degrees of the last forward's graph (the backward's)
};

namespace {

int mpnn_dims(const tlfea_mpnn* net, int blk, int* rows, int* cols) {
    if (blk == 0) {
        *rows = MW;
        *cols = MFEAT;
    } else if (blk <= net->layers) {
        *rows = MW;
        *cols = 2 * MW;
    } else {
        *rows = MOUT;
        *cols = MW;
    }
    return 0;
}

template <typename T>
int malloc_n(T** p, size_t n) {
    if (n == 0) n = 1;
    if (hipMalloc((void**)p, sizeof(T) * n) != hipSuccess) {
        set_error("device allocation of %zu bytes failed", sizeof(T) * n);
        return TLFEA_EHIP;
    }
    return TLFEA_OK;
}

void free_ws(tlfea_mpnn* net) {
    void* ps[] = {net->d_x, net->d_h, net->d_mg, net->d_dp16, net->d_dout16, net->d_pre, net->d_out,
                  net->d_dh, net->d_dm, net->d_dp, net->d_dout, net->d_part, net->d_lpart, net->d_deg};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    net->d_x = net->d_h = net->d_mg = net->d_dp16 = net->d_dout16 = nullptr;
    net->d_pre = net->d_out = net->d_dh = net->d_dm = net->d_dp = net->d_dout = nullptr;
    net->d_part = nullptr;
    net->d_deg = nullptr;
    net->d_lpart = nullptr;
    net->cap = 0;
}

int ensure_ws(tlfea_mpnn* net, int n) {
    if (n <= net->cap) return TLFEA_OK;
    free_ws(net);
    const size_t L = net->layers;
    const size_t nn = (size_t)n;
    const size_t nblk = std::min<size_t>((nn + WG_ROWS - 1) / WG_ROWS, WG_BLOCKS);
    int rc;
    if ((rc = malloc_n(&net->d_x, nn * MFEAT)) || (rc = malloc_n(&net->d_h, (L + 1) * nn * MW)) ||
        (rc = malloc_n(&net->d_mg, L * nn * MW)) || (rc = malloc_n(&net->d_dp16, nn * MW)) ||
        (rc = malloc_n(&net->d_dout16, nn * MOUT)) || (rc = malloc_n(&net->d_pre, (L + 1) * nn * MW)) ||
        (rc = malloc_n(&net->d_out, nn * MOUT)) || (rc = malloc_n(&net->d_dh, nn * MW)) ||
        (rc = malloc_n(&net->d_dm, nn * MW)) || (rc = malloc_n(&net->d_dp, nn * MW)) ||
        (rc = malloc_n(&net->d_dout, nn * MOUT)) || (rc = malloc_n(&net->d_part, nblk * MW * 2 * MW)) ||
        (rc = malloc_n(&net->d_lpart, (nn + 255) / 256 + 1)) || (rc = malloc_n(&net->d_deg, nn)))
        return rc;
    net->cap = n;
    return TLFEA_OK;
}

int refresh_bf16(tlfea_mpnn* net, hipStream_t s) {
    k_to_bf16<<<grid_for(net->n_params, 256), 256, 0, s>>>(net->n_params, net->d_w, net->d_w16);
    for (int blk = 0; blk < net->layers + 2; blk++) {
        int r, c;
        mpnn_dims(net, blk, &r, &c);
        k_to_bf16_t<<<grid_for(r * c, 256), 256, 0, s>>>(r, c, net->d_w + net->off_W[blk],
                                                         net->d_wt16 + net->off_W[blk]);
    }
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

int gemm_blocks(int M) { return (M + 63) / 64; }

// forward from the features in d_x; out -> d_out [n, 16]
int forward(tlfea_mpnn* net, int n, const int* off, const int* col, hipStream_t s) {
    const int L = net->layers;
    const size_t nn = (size_t)n;
    bf16* H = net->d_h;
    float* PRE = net->d_pre;
    // encoder
    k_gemm<MW><<<gemm_blocks(n), 256, 0, s>>>(n, net->d_x, MFEAT, net->d_w16 + net->off_W[0], MFEAT, nullptr, 0,
                                              nullptr, 0, net->d_w + net->off_b[0], 1, nullptr, PRE, H, nullptr, 0);
    for (int l = 1; l <= L; l++) {
        const bf16* h = H + (l - 1) * nn * MW;
        bf16* mg = net->d_mg + (l - 1) * nn * MW;
        k_agg<<<(n + 7) / 8, 256, 0, s>>>(n, off, col, h, mg);
        const bf16* W = net->d_w16 + net->off_W[l];  // [64 x 128]
        k_gemm<MW><<<gemm_blocks(n), 256, 0, s>>>(n, h, MW, W, 2 * MW, mg, MW, W + MW, 2 * MW,
                                                  net->d_w + net->off_b[l], 1, h, PRE + l * nn * MW,
                                                  H + l * nn * MW, nullptr, 0);
    }
    const int lo = L + 1;
    k_gemm<MOUT><<<gemm_blocks(n), 256, 0, s>>>(n, H + L * nn * MW, MW, net->d_w16 + net->off_W[lo], MW, nullptr, 0,
                                                nullptr, 0, net->d_w + net->off_b[lo], 0, nullptr, nullptr, nullptr,
                                                net->d_out, MOUT);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

// g (an N x K block with row stride ldg) += dY^T A
template <int N, int K>
void wgrad(tlfea_mpnn* net, int n, const bf16* dY, const bf16* A, float* g, int ldg, hipStream_t s) {
    const int nb = std::min((n + WG_ROWS - 1) / WG_ROWS, WG_BLOCKS);
    k_wgrad<N, K><<<nb, 256, 0, s>>>(n, dY, A, net->d_part);
    k_sum_partials<<<grid_for(N * K, 256), 256, 0, s>>>(nb, N * K, K, ldg, net->d_part, g);
}

void bgrad(tlfea_mpnn* net, int n, int N, const float* dY, float* g, hipStream_t s) {
    const int nb = std::min((n + WG_ROWS - 1) / WG_ROWS, WG_BLOCKS);
    k_colsum_partials<<<nb, 64, 0, s>>>(n, N, dY, net->d_part);
    k_sum_partials<<<1, 64, 0, s>>>(nb, N, N, N, net->d_part, g);
}

// backward of the loss (d_dout / d_dout16 set) through the network; gradients add to d_g
int backward(tlfea_mpnn* net, int n, const int* off, const int* col, hipStream_t s) {
    const int L = net->layers;
    const size_t nn = (size_t)n;
    const int lo = L + 1;
    const int64_t len = (int64_t)nn * MW;
    k_deg<<<grid_for(n, 256), 256, 0, s>>>(n, off, net->d_deg);
    // decoder: dW_out += dout^T h_L, db_out += colsum(dout), dh = dout W_out
    wgrad<MOUT, MW>(net, n, net->d_dout16, net->d_h + L * nn * MW, net->d_g + net->off_W[lo], MW, s);
    bgrad(net, n, MOUT, net->d_dout, net->d_g + net->off_b[lo], s);
    k_gemm<MW><<<gemm_blocks(n), 256, 0, s>>>(n, net->d_dout16, MOUT, net->d_wt16 + net->off_W[lo], MOUT, nullptr, 0,
                                              nullptr, 0, nullptr, 0, nullptr, nullptr, nullptr, net->d_dh, MW);
    for (int l = L; l >= 1; l--) {
        const bf16* h = net->d_h + (l - 1) * nn * MW;
        const bf16* mg = net->d_mg + (l - 1) * nn * MW;
        k_relu_bwd<<<grid_for(len, 256), 256, 0, s>>>(len, net->d_dh, net->d_pre + l * nn * MW, net->d_dp,
                                                      net->d_dp16);
        // dW_l = dpre^T [h, m]: columns 0..63 from h, 64..127 from m
        float* gW = net->d_g + net->off_W[l];
        wgrad<MW, MW>(net, n, net->d_dp16, h, gW, 2 * MW, s);
        wgrad<MW, MW>(net, n, net->d_dp16, mg, gW + MW, 2 * MW, s);
        bgrad(net, n, MW, net->d_dp, net->d_g + net->off_b[l], s);
        // d[h, m] = dpre W_l; W_l^T rows 0..63 give dh_direct, rows 64..127 give dm
        const bf16* WT = net->d_wt16 + net->off_W[l];  // [128 x 64]
        k_gemm<MW><<<gemm_blocks(n), 256, 0, s>>>(n, net->d_dp16, MW, WT + MW * MW, MW, nullptr, 0, nullptr, 0,
                                                  nullptr, 0, nullptr, nullptr, nullptr, net->d_dm, MW);
        k_gemm<MW><<<gemm_blocks(n), 256, 0, s>>>(n, net->d_dp16, MW, WT, MW, nullptr, 0, nullptr, 0, nullptr, 0,
                                                  nullptr, nullptr, nullptr, net->d_dp, MW);
        // dh_prev = dh (residual path) + dh_direct + agg^T(dm)
        k_add_f32<<<grid_for(len, 256), 256, 0, s>>>(len, net->d_dp, net->d_dh);
        k_agg_t<<<(n + 7) / 8, 256, 0, s>>>(n, off, col, net->d_deg, net->d_dm, net->d_dh);
    }
    // encoder
    k_relu_bwd<<<grid_for(len, 256), 256, 0, s>>>(len, net->d_dh, net->d_pre, net->d_dp, net->d_dp16);
    wgrad<MW, MFEAT>(net, n, net->d_dp16, net->d_x, net->d_g + net->off_W[0], MFEAT, s);
    bgrad(net, n, MW, net->d_dp, net->d_g + net->off_b[0], s);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

}  // namespace

extern "C" int tlfea_mpnn_create(int layers, tlfea_mpnn** out) {
    TLFEA_REQUIRE(out != nullptr, "null output handle");
    *out = nullptr;
    TLFEA_REQUIRE(layers >= 1 && layers <= 16, "layers out of range [1, 16]");
    auto* net = new tlfea_mpnn();
    net->layers = layers;
    int off = 0;
    for (int blk = 0; blk < layers + 2; blk++) {
        int r, c;
        mpnn_dims(net, blk, &r, &c);
        net->off_W.push_back(off);
        off += r * c;
        net->off_b.push_back(off);
        off += r;
    }
    net->n_params = off;
    std::vector<unsigned char> frozen(off, 0);
    {
        const int lo = layers + 1;
        for (int r = 3; r < MOUT; r++) {
            for (int c = 0; c < MW; c++) frozen[net->off_W[lo] + r * MW + c] = 1;
            frozen[net->off_b[lo] + r] = 1;
        }
    }
    int rc;
    if ((rc = malloc_n(&net->d_w, off)) || (rc = malloc_n(&net->d_g, off)) || (rc = malloc_n(&net->d_m1, off)) ||
        (rc = malloc_n(&net->d_m2, off)) || (rc = malloc_n(&net->d_frozen, off)) ||
        (rc = malloc_n(&net->d_w16, off)) || (rc = malloc_n(&net->d_wt16, off))) {
        tlfea_mpnn_destroy(net);
        return rc;
    }
    if (hipMemset(net->d_w, 0, sizeof(float) * off) || hipMemset(net->d_g, 0, sizeof(float) * off) ||
        hipMemset(net->d_m1, 0, sizeof(float) * off) || hipMemset(net->d_m2, 0, sizeof(float) * off) ||
        hipMemcpy(net->d_frozen, frozen.data(), off, hipMemcpyHostToDevice)) {
        set_error("network state initialisation failed");
        tlfea_mpnn_destroy(net);
        return TLFEA_EHIP;
    }
    if ((rc = refresh_bf16(net, nullptr))) {
        tlfea_mpnn_destroy(net);
        return rc;
    }
    *out = net;
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_destroy(tlfea_mpnn* net) {
    if (!net) return TLFEA_OK;
    free_ws(net);
    void* ps[] = {net->d_w, net->d_g, net->d_m1, net->d_m2, net->d_frozen, net->d_w16, net->d_wt16};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete net;
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_num_params(const tlfea_mpnn* net, int* n) {
    TLFEA_REQUIRE(net && n, "null argument");
    *n = net->n_params;
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_set_params(tlfea_mpnn* net, const float* host) {
    TLFEA_REQUIRE(net && host, "null argument");
    std::vector<float> w(host, host + net->n_params);
    const int lo = net->layers + 1;
    for (int r = 3; r < MOUT; r++) {  // the padded output rows stay zero
        for (int c = 0; c < MW; c++) w[net->off_W[lo] + r * MW + c] = 0.f;
        w[net->off_b[lo] + r] = 0.f;
    }
    TLFEA_HIP_TRY(hipMemcpy(net->d_w, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    TLFEA_HIP_TRY(hipMemset(net->d_m1, 0, sizeof(float) * net->n_params));
    TLFEA_HIP_TRY(hipMemset(net->d_m2, 0, sizeof(float) * net->n_params));
    net->adam_t = 0;
    if (int rc = refresh_bf16(net, nullptr)) return rc;
    TLFEA_HIP_TRY(hipDeviceSynchronize());
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_get_params(const tlfea_mpnn* net, float* host) {
    TLFEA_REQUIRE(net && host, "null argument");
    TLFEA_HIP_TRY(hipMemcpy(host, net->d_w, sizeof(float) * net->n_params, hipMemcpyDeviceToHost));
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_get_grads(const tlfea_mpnn* net, float* host) {
    TLFEA_REQUIRE(net && host, "null argument");
    TLFEA_HIP_TRY(hipMemcpy(host, net->d_g, sizeof(float) * net->n_params, hipMemcpyDeviceToHost));
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_zero_grads(tlfea_mpnn* net, void* stream) {
    TLFEA_REQUIRE(net != nullptr, "null network");
    TLFEA_HIP_TRY(hipMemsetAsync(net->d_g, 0, sizeof(float) * net->n_params, as_stream(stream)));
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_forward(tlfea_mpnn* net, int n, const unsigned short* feat, const int* off,
                                  const int* col, float* out, void* stream) {
    TLFEA_REQUIRE(net && feat && off && col && out && n > 0, "bad argument");
    hipStream_t s = as_stream(stream);
    if (int rc = ensure_ws(net, n)) return rc;
    TLFEA_HIP_TRY(hipMemcpyAsync(net->d_x, feat, sizeof(bf16) * (size_t)n * MFEAT, hipMemcpyDeviceToDevice, s));
    if (int rc = forward(net, n, off, col, s)) return rc;
    TLFEA_HIP_TRY(hipMemcpyAsync(out, net->d_out, sizeof(float) * (size_t)n * MOUT, hipMemcpyDeviceToDevice, s));
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_loss_grad(tlfea_mpnn* net, int n, const unsigned short* feat, const int* off,
                                    const int* col, const float* target, const float* weight, double* loss,
                                    void* stream) {
    TLFEA_REQUIRE(net && feat && off && col && target && weight && n > 0, "bad argument");
    hipStream_t s = as_stream(stream);
    if (int rc = ensure_ws(net, n)) return rc;
    // the weight sum fixes the loss normalisation (host reduction of a small vector)
    std::vector<float> w(n);
    TLFEA_HIP_TRY(hipMemcpyAsync(w.data(), weight, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    TLFEA_HIP_TRY(hipStreamSynchronize(s));
    double wsum = 0.0;
    for (float v : w) wsum += v;
    TLFEA_REQUIRE(wsum > 0.0, "all loss weights are zero");
    TLFEA_HIP_TRY(hipMemcpyAsync(net->d_x, feat, sizeof(bf16) * (size_t)n * MFEAT, hipMemcpyDeviceToDevice, s));
    if (int rc = forward(net, n, off, col, s)) return rc;
    const int nb = (n + 255) / 256;
    k_loss<<<nb, 256, 0, s>>>(n, net->d_out, target, weight, (float)(1.0 / wsum), net->d_dout16, net->d_dout,
                              net->d_lpart);
    TLFEA_LAUNCH_CHECK();
    if (int rc = backward(net, n, off, col, s)) return rc;
    if (loss) {
        std::vector<double> lp(nb);
        TLFEA_HIP_TRY(hipMemcpyAsync(lp.data(), net->d_lpart, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        TLFEA_HIP_TRY(hipStreamSynchronize(s));
        double l = 0.0;
        for (double v : lp) l += v;
        *loss = l;
    }
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_adam(tlfea_mpnn* net, double lr, double beta1, double beta2, double eps, void* stream) {
    TLFEA_REQUIRE(net != nullptr, "null network");
    hipStream_t s = as_stream(stream);
    net->adam_t += 1;
    const double c1 = 1.0 / (1.0 - std::pow(beta1, net->adam_t)), c2 = 1.0 / (1.0 - std::pow(beta2, net->adam_t));
    k_adam<<<grid_for(net->n_params, 256), 256, 0, s>>>(net->n_params, net->d_w, net->d_g, net->d_m1, net->d_m2,
                                                        net->d_frozen, (float)lr, (float)beta1, (float)beta2,
                                                        (float)eps, (float)c1, (float)c2);
    TLFEA_LAUNCH_CHECK();
    return refresh_bf16(net, s);
}

extern "C" int tlfea_mpnn_features(int n, const int* node_inst, const double* x, const double* y, const double* z,
                                   const unsigned char* fixed_dof_mask, const double* f, const double* inst_par,
                                   unsigned short* feat, void* stream) {
    TLFEA_REQUIRE(n > 0 && node_inst && x && y && z && fixed_dof_mask && f && inst_par && feat, "bad argument");
    k_features<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(n, node_inst, x, y, z, fixed_dof_mask, f, inst_par,
                                                               feat);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_scale_output(int n, const int* node_inst, const float* out, const double* scale,
                                       const unsigned char* fixed_dof_mask, double* u, void* stream) {
    TLFEA_REQUIRE(n > 0 && node_inst && out && scale && fixed_dof_mask && u, "bad argument");
    k_scale_out<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(n, node_inst, out, scale, fixed_dof_mask, u);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}

extern "C" int tlfea_mpnn_backward(tlfea_mpnn* net, int n, const int* off, const int* col, const float* dout,
                                   void* stream) {
    TLFEA_REQUIRE(net && off && col && dout && n > 0, "bad argument");
    TLFEA_REQUIRE(n <= net->cap, "backward needs the forward of the same nodes first");
    hipStream_t s = as_stream(stream);
    k_set_dout<<<grid_for(n, 256), 256, 0, s>>>(n, dout, net->d_dout, net->d_dout16);
    TLFEA_LAUNCH_CHECK();
    return backward(net, n, off, col, s);
}

extern "C" int tlfea_mpnn_energy_grad(int n, const int* node_inst, const double* scale, const double* Ke,
                                      const double* w_over_E, float* dout, void* stream) {
    TLFEA_REQUIRE(n > 0 && node_inst && scale && Ke && w_over_E && dout, "bad argument");
    k_energy_dout<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(n, node_inst, scale, Ke, w_over_E, dout);
    TLFEA_LAUNCH_CHECK();
    return TLFEA_OK;
}
