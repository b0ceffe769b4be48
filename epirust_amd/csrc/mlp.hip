// Surrogate MLP kernels for gfx950: bf16 operands on MFMA with FP32 accumulation.
//
// GEMM: C[M x N] = A[M x K] * B[K x N] with arbitrary element strides for A and B (so
// one kernel serves X*W, X^T*dZ and dZ*W^T), a 64x64 block tile, BK = 32, four 64-lane
// waves each owning a 32x32 sub-tile computed with v_mfma_f32_32x32x16_bf16 (gfx950; A
// and B staged through LDS as 16-byte vectors, two buffers, the next tile prefetched into
// registers during the current one's MFMAs), and a fused epilogue: + bias[n], optional
// GELU (tanh form), optional FP32 and/or bf16 stores, optional store of the
// pre-activation.
//
// v_mfma_f32_32x32x16_bf16 lane layout (64 lanes, wave64):
//   A (32x16): lane l holds row l%32, k = 8*(l/32) .. +7 (eight bf16 in 128 bits)
//   B (16x32): lane l holds col l%32, k = 8*(l/32) .. +7
//   D (32x32, 16 floats / lane): lane l, element r -> row 8*(r/4) + 4*(l/32) + r%4, col l%32
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mlp.hpp"

namespace epi {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int kLdsK = kBK + 8;  // LDS row: kBK k of one m (or n) plus 8 of padding
constexpr int kChunks = kBK / 8;  // 16-byte vectors per tile row along k
constexpr int kVecs = kBM * kBK / (256 * 8);  // 16-byte vectors per thread per operand tile

// How one 64 x kBK operand tile (rows r = m or n, columns k) is read from global memory.
enum : int { TILE_K_CONTIG = 0, TILE_R_CONTIG = 1, TILE_ELEMENT = 2 };

// Element (r, k) of an operand is X[r * ld_r + k * ld_k]. A whole tile inside the matrix
// whose contiguous dimension is 16-byte aligned is read as 16-byte vectors (kVecs of 8
// bf16 per thread, 256 threads): along k (ld_k == 1) or along r (ld_r == 1). Otherwise, at
// ragged edges or odd strides, element by element with bounds checks.
__device__ __forceinline__ int tile_mode(const uint16_t* X, int64_t ld_r, int64_t ld_k, int r0, int k0, int R,
                                         int K) {
  const bool full = r0 + 64 <= R && k0 + kBK <= K;
  const bool aligned = (reinterpret_cast<uintptr_t>(X) & 15u) == 0;
  if (full && aligned && ld_k == 1 && (ld_r & 7) == 0) return TILE_K_CONTIG;
  if (full && aligned && ld_r == 1 && (ld_k & 7) == 0) return TILE_R_CONTIG;
  return TILE_ELEMENT;
}

// Vector e (e = tid + 256 i, i < kVecs) of a tile: along k, row e / kChunks with k
// (e % kChunks) * 8 .. +7; along r, k e / 8 with rows (e % 8) * 8 .. +7.
struct TileRegs {
  u16x8 v[kVecs];
};

__device__ __forceinline__ void load_tile(TileRegs& t, const uint16_t* X, int64_t ld_r, int64_t ld_k, int r0,
                                          int k0, int R, int K, int mode, int tid) {
  if (mode == TILE_K_CONTIG) {
#pragma unroll
    for (int i = 0; i < kVecs; ++i) {
      const int e = tid + 256 * i;
      t.v[i] = *reinterpret_cast<const u16x8*>(X + static_cast<int64_t>(r0 + e / kChunks) * ld_r + k0 + (e % kChunks) * 8);
    }
  } else if (mode == TILE_R_CONTIG) {
#pragma unroll
    for (int i = 0; i < kVecs; ++i) {
      const int e = tid + 256 * i;
      t.v[i] = *reinterpret_cast<const u16x8*>(X + static_cast<int64_t>(k0 + (e >> 3)) * ld_k + r0 + (e & 7) * 8);
    }
  } else {  // ragged edges and odd strides: the along-k assignment, element by element
#pragma unroll 1
    for (int i = 0; i < kVecs; ++i) {
      const int e = tid + 256 * i;
      const int r = r0 + e / kChunks, kb = k0 + (e % kChunks) * 8;
      u16x8 v;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = kb + q;
        v[q] = (r < R && k < K) ? X[static_cast<int64_t>(r) * ld_r + static_cast<int64_t>(k) * ld_k] : 0;
      }
      t.v[i] = v;
    }
  }
}

// LDS tile s[r][k] (k contiguous): 16-byte stores, or 2-byte stores for a tile read
// along r.
__device__ __forceinline__ void store_tile(uint16_t (*s)[kLdsK], const TileRegs& t, int mode, int tid) {
#pragma unroll
  for (int i = 0; i < kVecs; ++i) {
    const int e = tid + 256 * i;
    if (mode == TILE_R_CONTIG) {
      const int k = e >> 3, rb = (e & 7) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) s[rb + q][k] = t.v[i][q];
    } else {
      *reinterpret_cast<u16x8*>(&s[e / kChunks][(e % kChunks) * 8]) = t.v[i];
    }
  }
}

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

// C tile 64 x 64 per block of four wave64s (2 x 2), each a 32 x 32 sub-tile accumulated
// with v_mfma_f32_32x32x16_bf16 (gfx950): per kBK-wide k step, kBK / 16 MFMAs whose A operand is
// row (lane % 32) with k 8 * (lane / 32) .. +7, read from LDS as one 16-byte vector, and B
// likewise by column. The next k tile's global loads are issued before the current one's
// MFMAs (register prefetch), into the other of two LDS buffers, so each k step has one
// barrier.
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __align__(16) uint16_t sA[2][kBM][kLdsK];
  __shared__ __align__(16) uint16_t sB[2][kBN][kLdsK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc = {0};

  int ma = tile_mode(g.A, g.lda_m, g.lda_k, m0, 0, g.M, g.K);
  int mb = tile_mode(g.B, g.ldb_n, g.ldb_k, n0, 0, g.N, g.K);
  TileRegs va, vb;
  load_tile(va, g.A, g.lda_m, g.lda_k, m0, 0, g.M, g.K, ma, tid);
  load_tile(vb, g.B, g.ldb_n, g.ldb_k, n0, 0, g.N, g.K, mb, tid);
  int buf = 0;
  for (int k0 = 0; k0 < g.K; k0 += kBK) {
    store_tile(sA[buf], va, ma, tid);
    store_tile(sB[buf], vb, mb, tid);
    __syncthreads();
    const int k1 = k0 + kBK;
    if (k1 < g.K) {  // uniform: prefetch the next tile while this one is multiplied
      ma = tile_mode(g.A, g.lda_m, g.lda_k, m0, k1, g.M, g.K);
      mb = tile_mode(g.B, g.ldb_n, g.ldb_k, n0, k1, g.N, g.K);
      load_tile(va, g.A, g.lda_m, g.lda_k, m0, k1, g.M, g.K, ma, tid);
      load_tile(vb, g.B, g.ldb_n, g.ldb_k, n0, k1, g.N, g.K, mb, tid);
    }
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      const int kq = kk + 8 * (lane >> 5);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[buf][wm + (lane & 31)][kq]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[buf][wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    buf ^= 1;
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
