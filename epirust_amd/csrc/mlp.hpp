// Surrogate MLP: launch interface of the bf16-MFMA kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace epi {

enum : int { ACT_NONE = 0, ACT_GELU = 1 };

// C[M x N] = alpha * A * B (+ bias[n]) (+ C32 if beta_accumulate), then the activation.
// A element (m, k) is A[m * lda_m + k * lda_k]; B element (k, n) is B[k * ldb_k + n * ldb_n].
struct GemmArgs {
  const uint16_t* A;
  const uint16_t* B;
  int64_t lda_m, lda_k, ldb_k, ldb_n;
  int M, N, K;
  float alpha;
  const float* bias;
  int act;
  int beta_accumulate;
  float* C32;         // optional fp32 output (row pitch ldc)
  uint16_t* C16;      // optional bf16 output (row pitch ldc)
  float* preact;      // optional fp32 pre-activation (row pitch ldc)
  int64_t ldc;
};

hipError_t mlp_gemm(const GemmArgs& g, hipStream_t s);
hipError_t mlp_gelu_backward(const float* dA, const float* Z, float* dZ32, uint16_t* dZ16, int64_t n, hipStream_t s);
hipError_t mlp_mse(const float* Y, const float* T, float* dY32, uint16_t* dY16, float* loss_parts, int64_t n,
                   hipStream_t s, uint32_t* n_parts);
hipError_t mlp_colsum(const float* dZ, float* out, int M, int N, hipStream_t s);
hipError_t mlp_adam(float* w, uint16_t* w16, const float* grad, float* m, float* v, int64_t n, float lr, float b1,
                    float b2, float eps, int step, hipStream_t s);
hipError_t mlp_to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t s);

}  // namespace epi
