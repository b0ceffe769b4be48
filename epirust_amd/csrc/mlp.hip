// Surrogate MLP kernels for gfx950: bf16 operands on MFMA with FP32 accumulation.
//
// GEMM: C[M x N] = A[M x K] * B[K x N] with arbitrary element strides for A and B (so
// one kernel serves X*W, X^T*dZ and dZ*W^T), a 64x64 block tile, BK = 32, four 64-lane
// waves each owning a 32x32 sub-tile computed with v_mfma_f32_32x32x8_bf16 (A and B
// staged through LDS), and a fused epilogue: + bias[n], optional GELU (tanh form),
// optional FP32 and/or bf16 stores, optional store of the pre-activation.
//
// v_mfma_f32_32x32x8_bf16 lane layout (64 lanes, wave64):
//   A (32x8): lane l holds row l%32, k = 4*(l/32) .. +3 (four bf16 in 64 bits)
//   B (8x32): lane l holds col l%32, k = 4*(l/32) .. +3
//   D (32x32, 16 floats / lane): lane l, element r -> row 8*(r/4) + 4*(l/32) + r%4, col l%32
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mlp.hpp"

namespace epi {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 64, kBN = 64, kBK = 32;

// fp32 -> bf16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return static_cast<uint16_t>((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
__device__ __forceinline__ float gelu(float x) {
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float x2 = x * x;
  const float u = 0.7978845608028654f * (x + 0.044715f * x2 * x);
  const float t = tanhf(u);
  const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x2);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ uint16_t sA[kBK][kBM + 4];  // k-major: sA[k][m]
  __shared__ uint16_t sB[kBK][kBN + 4];  // sB[k][n]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc = {0};

  for (int k0 = 0; k0 < g.K; k0 += kBK) {
    // stage A (64 x 32) and B (32 x 64): 2048 elements each, 8 per thread
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + r * 256;
      const int am = e % kBM, ak = e / kBM;  // consecutive threads -> consecutive m
      const int gm = m0 + am, gk = k0 + ak;
      sA[ak][am] = (gm < g.M && gk < g.K) ? g.A[static_cast<int64_t>(gm) * g.lda_m + static_cast<int64_t>(gk) * g.lda_k] : 0;
      const int bn = e % kBN, bk = e / kBN;
      const int gn = n0 + bn, gk2 = k0 + bk;
      sB[bk][bn] = (gn < g.N && gk2 < g.K) ? g.B[static_cast<int64_t>(gk2) * g.ldb_k + static_cast<int64_t>(gn) * g.ldb_n] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 8) {
      const int kb = kk + 4 * (lane >> 5);
      bf16x4 a, b;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = static_cast<short>(sA[kb + q][wm + (lane & 31)]);
        b[q] = static_cast<short>(sB[kb + q][wn + (lane & 31)]);
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue
  const int col = n0 + wn + (lane & 31);
  if (col >= g.N) return;
  const float bias = g.bias ? g.bias[col] : 0.0f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (row >= g.M) continue;
    float z = g.alpha * acc[r] + bias;
    const int64_t o = static_cast<int64_t>(row) * g.ldc + col;
    if (g.preact) g.preact[o] = z;
    if (g.act == ACT_GELU) z = gelu(z);
    if (g.beta_accumulate && g.C32) z += g.C32[o];
    if (g.C32) g.C32[o] = z;
    if (g.C16) g.C16[o] = to_bf16(z);
  }
}

// dZ = dA * gelu'(Z), written as fp32 and bf16
__global__ void k_gelu_backward(const float* dA, const float* Z, float* dZ32, uint16_t* dZ16, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = dA[i] * gelu_grad(Z[i]);
  dZ32[i] = v;
  dZ16[i] = to_bf16(v);
}

// MSE over all outputs: loss = mean((Y - T)^2); dY = 2 (Y - T) / count
__global__ void k_mse(const float* Y, const float* T, float* dY32, uint16_t* dY16, float* loss_parts, int64_t n,
                      float inv_count) {
  __shared__ float part[4];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float d = 0.0f;
  if (i < n) {
    d = Y[i] - T[i];
    const float g = 2.0f * d * inv_count;
    dY32[i] = g;
    dY16[i] = to_bf16(g);
  }
  float s = d * d;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += part[w];
    loss_parts[blockIdx.x] = t * inv_count;
  }
}

// column sums of dZ (M x N): the bias gradient
__global__ void k_colsum(const float* dZ, float* out, int M, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += dZ[static_cast<int64_t>(m) * N + n];
  out[n] = s;
}

// Adam (Kingma & Ba): fp32 master weights, bf16 copy for the GEMMs
__global__ void k_adam(float* w, uint16_t* w16, const float* grad, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  const float nw = w[i] - upd;
  w[i] = nw;
  if (w16) w16[i] = to_bf16(nw);
}

__global__ void k_to_bf16(const float* x, uint16_t* y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}

inline dim3 grid1(int64_t n, int b) { return dim3(static_cast<uint32_t>((n + b - 1) / b)); }

}  // namespace

hipError_t mlp_gemm(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipSuccess;
  dim3 grid((g.N + kBN - 1) / kBN, (g.M + kBM - 1) / kBM);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t mlp_gelu_backward(const float* dA, const float* Z, float* dZ32, uint16_t* dZ16, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gelu_backward, grid1(n, 256), dim3(256), 0, s, dA, Z, dZ32, dZ16, n);
  return hipGetLastError();
}

hipError_t mlp_mse(const float* Y, const float* T, float* dY32, uint16_t* dY16, float* loss_parts, int64_t n,
                   hipStream_t s, uint32_t* n_parts) {
  *n_parts = static_cast<uint32_t>((n + 255) / 256);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mse, grid1(n, 256), dim3(256), 0, s, Y, T, dY32, dY16, loss_parts, n,
                     1.0f / static_cast<float>(n));
  return hipGetLastError();
}

hipError_t mlp_colsum(const float* dZ, float* out, int M, int N, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, grid1(N, 256), dim3(256), 0, s, dZ, out, M, N);
  return hipGetLastError();
}

hipError_t mlp_adam(float* w, uint16_t* w16, const float* grad, float* m, float* v, int64_t n, float lr, float b1,
                    float b2, float eps, int step, hipStream_t s) {
  const float c1 = 1.0f - std::pow(b1, static_cast<float>(step));
  const float c2 = 1.0f - std::pow(b2, static_cast<float>(step));
  hipLaunchKernelGGL(k_adam, grid1(n, 256), dim3(256), 0, s, w, w16, grad, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16, grid1(n, 256), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}

}  // namespace epi
