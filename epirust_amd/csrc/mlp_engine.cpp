// Host runtime of the surrogate MLP (C ABI in include/epirust_gpu.h): widths
// w0 -> w1 -> ... -> wL, GELU after every hidden layer, identity output, MSE loss, Adam.
// Parameters are fp32 masters with bf16 copies; activations are bf16; every GEMM runs on
// the bf16 MFMA kernel with fp32 accumulation.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "epirust_gpu.h"
#include "mlp.hpp"

using namespace epi;

namespace {

thread_local std::string g_mlp_error;

int mfail(int code, const std::string& m) {
  g_mlp_error = m;
  return code;
}

#define MHIP(expr)                                                                                          \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) return mfail(EPI_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

uint16_t host_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return static_cast<uint16_t>((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

struct Buf {
  std::vector<void*> owned;
  template <class T>
  hipError_t get(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
      owned.push_back(q);
      (void)hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T));
    }
    *p = static_cast<T*>(q);
    return e;
  }
  ~Buf() {
    for (void* p : owned) (void)hipFree(p);
  }
};

}  // namespace

struct epi_mlp {
  std::vector<int> w;  // widths, L+1
  int L = 0, max_batch = 0, step = 0;
  hipStream_t stream = nullptr;
  Buf mem;
  std::vector<float*> W, b, gW, gb, mW, vW, mb, vb, Z, dA, dZ32;
  std::vector<uint16_t*> W16, act16, dZ16;  // act16[0] = input, act16[l+1] = output of layer l
  float* Y = nullptr;
  float* T = nullptr;
  float* loss_parts = nullptr;
  ~epi_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  size_t param_count() const {
    size_t n = 0;
    for (int l = 0; l < L; ++l) n += static_cast<size_t>(w[l]) * w[l + 1] + w[l + 1];
    return n;
  }
};

namespace {

int forward(epi_mlp* m, int B) {
  for (int l = 0; l < m->L; ++l) {
    const int K = m->w[l], N = m->w[l + 1];
    const bool last = l == m->L - 1;
    GemmArgs g{};
    g.A = m->act16[l];
    g.lda_m = K;
    g.lda_k = 1;
    g.B = m->W16[l];
    g.ldb_k = N;
    g.ldb_n = 1;
    g.M = B;
    g.N = N;
    g.K = K;
    g.alpha = 1.0f;
    g.bias = m->b[l];
    g.act = last ? ACT_NONE : ACT_GELU;
    g.preact = last ? nullptr : m->Z[l];
    g.C16 = last ? nullptr : m->act16[l + 1];
    g.C32 = last ? m->Y : nullptr;
    g.ldc = N;
    MHIP(mlp_gemm(g, m->stream));
  }
  return EPI_OK;
}

int backward(epi_mlp* m, int B) {
  for (int l = m->L - 1; l >= 0; --l) {
    const int K = m->w[l], N = m->w[l + 1];
    // gW = act_l^T * dZ_l  (K x N)
    GemmArgs g{};
    g.A = m->act16[l];
    g.lda_m = 1;
    g.lda_k = K;
    g.B = m->dZ16[l];
    g.ldb_k = N;
    g.ldb_n = 1;
    g.M = K;
    g.N = N;
    g.K = B;
    g.alpha = 1.0f;
    g.act = ACT_NONE;
    g.C32 = m->gW[l];
    g.ldc = N;
    MHIP(mlp_gemm(g, m->stream));
    MHIP(mlp_colsum(m->dZ32[l], m->gb[l], B, N, m->stream));
    if (l > 0) {
      // dA_{l} = dZ_l * W_l^T  (B x K), then dZ_{l-1} = dA * gelu'(Z_{l-1})
      GemmArgs h{};
      h.A = m->dZ16[l];
      h.lda_m = N;
      h.lda_k = 1;
      h.B = m->W16[l];
      h.ldb_k = 1;
      h.ldb_n = N;
      h.M = B;
      h.N = K;
      h.K = N;
      h.alpha = 1.0f;
      h.act = ACT_NONE;
      h.C32 = m->dA[l];
      h.ldc = K;
      MHIP(mlp_gemm(h, m->stream));
      MHIP(mlp_gelu_backward(m->dA[l], m->Z[l - 1], m->dZ32[l - 1], m->dZ16[l - 1], static_cast<int64_t>(B) * K,
                             m->stream));
    }
  }
  return EPI_OK;
}

int load_xt(epi_mlp* m, const float* x, const float* t, int B) {
  if (B <= 0 || B > m->max_batch) return mfail(EPI_ERR_ARG, "batch out of range");
  const size_t nx = static_cast<size_t>(B) * m->w[0];
  std::vector<uint16_t> x16(nx);
  for (size_t i = 0; i < nx; ++i) x16[i] = host_bf16(x[i]);
  MHIP(hipMemcpyAsync(m->act16[0], x16.data(), nx * 2, hipMemcpyHostToDevice, m->stream));
  if (t)
    MHIP(hipMemcpyAsync(m->T, t, static_cast<size_t>(B) * m->w[m->L] * 4, hipMemcpyHostToDevice, m->stream));
  MHIP(hipStreamSynchronize(m->stream));
  return EPI_OK;
}

int loss_and_grad(epi_mlp* m, int B, float* loss) {
  const int64_t n = static_cast<int64_t>(B) * m->w[m->L];
  uint32_t parts = 0;
  MHIP(mlp_mse(m->Y, m->T, m->dZ32[m->L - 1], m->dZ16[m->L - 1], m->loss_parts, n, m->stream, &parts));
  if (loss) {
    std::vector<float> lp(parts);
    MHIP(hipMemcpyAsync(lp.data(), m->loss_parts, parts * 4, hipMemcpyDeviceToHost, m->stream));
    MHIP(hipStreamSynchronize(m->stream));
    double s = 0;
    for (float v : lp) s += v;
    *loss = static_cast<float>(s);
  }
  return EPI_OK;
}

}  // namespace

extern "C" {

const char* epi_mlp_last_error(void) { return g_mlp_error.c_str(); }

int epi_mlp_create(const uint32_t* widths, uint32_t n_widths, uint64_t seed, uint32_t max_batch, epi_mlp** out) {
  if (!widths || n_widths < 2 || !out || max_batch == 0) return mfail(EPI_ERR_ARG, "bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return mfail(EPI_ERR_DEVICE, "no HIP device available");
  std::unique_ptr<epi_mlp> m(new epi_mlp());
  for (uint32_t i = 0; i < n_widths; ++i) {
    if (widths[i] == 0 || widths[i] > 65536) return mfail(EPI_ERR_ARG, "layer width out of range");
    m->w.push_back(static_cast<int>(widths[i]));
  }
  m->L = static_cast<int>(n_widths) - 1;
  m->max_batch = static_cast<int>(max_batch);
  MHIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  const int L = m->L, B = m->max_batch;
  m->W.resize(L); m->b.resize(L); m->gW.resize(L); m->gb.resize(L); m->mW.resize(L); m->vW.resize(L);
  m->mb.resize(L); m->vb.resize(L); m->Z.resize(L); m->dA.resize(L); m->dZ32.resize(L); m->W16.resize(L);
  m->dZ16.resize(L); m->act16.resize(L + 1);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < L; ++l) {
    const size_t K = m->w[l], N = m->w[l + 1];
    MHIP(m->mem.get(&m->W[l], K * N));
    MHIP(m->mem.get(&m->W16[l], K * N));
    MHIP(m->mem.get(&m->b[l], N));
    MHIP(m->mem.get(&m->gW[l], K * N));
    MHIP(m->mem.get(&m->gb[l], N));
    MHIP(m->mem.get(&m->mW[l], K * N));
    MHIP(m->mem.get(&m->vW[l], K * N));
    MHIP(m->mem.get(&m->mb[l], N));
    MHIP(m->mem.get(&m->vb[l], N));
    MHIP(m->mem.get(&m->Z[l], static_cast<size_t>(B) * N));
    MHIP(m->mem.get(&m->dA[l], static_cast<size_t>(B) * K));
    MHIP(m->mem.get(&m->dZ32[l], static_cast<size_t>(B) * N));
    MHIP(m->mem.get(&m->dZ16[l], static_cast<size_t>(B) * N));
    // He-normal initialisation for GELU layers, biases zero
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / static_cast<float>(K)));
    std::vector<float> w(K * N);
    for (auto& v : w) v = nd(rng);
    MHIP(hipMemcpy(m->W[l], w.data(), w.size() * 4, hipMemcpyHostToDevice));
    MHIP(mlp_to_bf16(m->W[l], m->W16[l], static_cast<int64_t>(K * N), m->stream));
  }
  for (int l = 0; l <= L; ++l) MHIP(m->mem.get(&m->act16[l], static_cast<size_t>(B) * m->w[l]));
  MHIP(m->mem.get(&m->Y, static_cast<size_t>(B) * m->w[L]));
  MHIP(m->mem.get(&m->T, static_cast<size_t>(B) * m->w[L]));
  MHIP(m->mem.get(&m->loss_parts, (static_cast<size_t>(B) * m->w[L] + 255) / 256));
  MHIP(hipStreamSynchronize(m->stream));
  *out = m.release();
  return EPI_OK;
}

void epi_mlp_destroy(epi_mlp* m) { delete m; }

uint64_t epi_mlp_param_count(epi_mlp* m) { return m ? m->param_count() : 0; }

int epi_mlp_get_params(epi_mlp* m, float* out, uint64_t n) {
  if (!m || !out || n != m->param_count()) return mfail(EPI_ERR_ARG, "parameter count mismatch");
  size_t o = 0;
  for (int l = 0; l < m->L; ++l) {
    const size_t kn = static_cast<size_t>(m->w[l]) * m->w[l + 1];
    MHIP(hipMemcpy(out + o, m->W[l], kn * 4, hipMemcpyDeviceToHost));
    o += kn;
    MHIP(hipMemcpy(out + o, m->b[l], m->w[l + 1] * 4, hipMemcpyDeviceToHost));
    o += m->w[l + 1];
  }
  return EPI_OK;
}

int epi_mlp_set_params(epi_mlp* m, const float* in, uint64_t n) {
  if (!m || !in || n != m->param_count()) return mfail(EPI_ERR_ARG, "parameter count mismatch");
  size_t o = 0;
  for (int l = 0; l < m->L; ++l) {
    const size_t kn = static_cast<size_t>(m->w[l]) * m->w[l + 1];
    MHIP(hipMemcpy(m->W[l], in + o, kn * 4, hipMemcpyHostToDevice));
    MHIP(mlp_to_bf16(m->W[l], m->W16[l], static_cast<int64_t>(kn), m->stream));
    o += kn;
    MHIP(hipMemcpy(m->b[l], in + o, m->w[l + 1] * 4, hipMemcpyHostToDevice));
    o += m->w[l + 1];
  }
  m->step = 0;
  for (int l = 0; l < m->L; ++l) {
    const size_t kn = static_cast<size_t>(m->w[l]) * m->w[l + 1];
    MHIP(hipMemsetAsync(m->mW[l], 0, kn * 4, m->stream));
    MHIP(hipMemsetAsync(m->vW[l], 0, kn * 4, m->stream));
    MHIP(hipMemsetAsync(m->mb[l], 0, m->w[l + 1] * 4, m->stream));
    MHIP(hipMemsetAsync(m->vb[l], 0, m->w[l + 1] * 4, m->stream));
  }
  MHIP(hipStreamSynchronize(m->stream));
  return EPI_OK;
}

int epi_mlp_forward(epi_mlp* m, const float* x, uint32_t batch, float* y) {
  if (!m || !x || !y) return mfail(EPI_ERR_ARG, "null argument");
  int rc = load_xt(m, x, nullptr, static_cast<int>(batch));
  if (rc) return rc;
  rc = forward(m, static_cast<int>(batch));
  if (rc) return rc;
  MHIP(hipMemcpyAsync(y, m->Y, static_cast<size_t>(batch) * m->w[m->L] * 4, hipMemcpyDeviceToHost, m->stream));
  MHIP(hipStreamSynchronize(m->stream));
  return EPI_OK;
}

int epi_mlp_gradients(epi_mlp* m, const float* x, const float* t, uint32_t batch, float* grads, float* loss) {
  if (!m || !x || !t || !grads) return mfail(EPI_ERR_ARG, "null argument");
  int rc = load_xt(m, x, t, static_cast<int>(batch));
  if (!rc) rc = forward(m, static_cast<int>(batch));
  if (!rc) rc = loss_and_grad(m, static_cast<int>(batch), loss);
  if (!rc) rc = backward(m, static_cast<int>(batch));
  if (rc) return rc;
  MHIP(hipStreamSynchronize(m->stream));
  size_t o = 0;
  for (int l = 0; l < m->L; ++l) {
    const size_t kn = static_cast<size_t>(m->w[l]) * m->w[l + 1];
    MHIP(hipMemcpy(grads + o, m->gW[l], kn * 4, hipMemcpyDeviceToHost));
    o += kn;
    MHIP(hipMemcpy(grads + o, m->gb[l], m->w[l + 1] * 4, hipMemcpyDeviceToHost));
    o += m->w[l + 1];
  }
  return EPI_OK;
}

int epi_mlp_train_step(epi_mlp* m, const float* x, const float* t, uint32_t batch, float lr, float* loss) {
  if (!m || !x || !t) return mfail(EPI_ERR_ARG, "null argument");
  int rc = load_xt(m, x, t, static_cast<int>(batch));
  if (!rc) rc = forward(m, static_cast<int>(batch));
  if (!rc) rc = loss_and_grad(m, static_cast<int>(batch), loss);
  if (!rc) rc = backward(m, static_cast<int>(batch));
  if (rc) return rc;
  ++m->step;
  for (int l = 0; l < m->L; ++l) {
    const int64_t kn = static_cast<int64_t>(m->w[l]) * m->w[l + 1];
    MHIP(mlp_adam(m->W[l], m->W16[l], m->gW[l], m->mW[l], m->vW[l], kn, lr, 0.9f, 0.999f, 1e-8f, m->step, m->stream));
    MHIP(mlp_adam(m->b[l], nullptr, m->gb[l], m->mb[l], m->vb[l], m->w[l + 1], lr, 0.9f, 0.999f, 1e-8f, m->step,
                  m->stream));
  }
  MHIP(hipStreamSynchronize(m->stream));
  return EPI_OK;
}

int epi_gemm_bf16(const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K, int trans_a, int trans_b) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0) return mfail(EPI_ERR_ARG, "bad arguments");
  uint16_t *dA = nullptr, *dB = nullptr;
  float* dC = nullptr;
  Buf mem;
  MHIP(mem.get(&dA, static_cast<size_t>(M) * K));
  MHIP(mem.get(&dB, static_cast<size_t>(K) * N));
  MHIP(mem.get(&dC, static_cast<size_t>(M) * N));
  MHIP(hipMemcpy(dA, A, static_cast<size_t>(M) * K * 2, hipMemcpyHostToDevice));
  MHIP(hipMemcpy(dB, B, static_cast<size_t>(K) * N * 2, hipMemcpyHostToDevice));
  GemmArgs g{};
  g.A = dA;
  g.B = dB;
  // A is stored M x K row-major unless trans_a (then K x M); likewise B
  g.lda_m = trans_a ? 1 : K;
  g.lda_k = trans_a ? M : 1;
  g.ldb_k = trans_b ? 1 : N;
  g.ldb_n = trans_b ? K : 1;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = 1.0f;
  g.act = ACT_NONE;
  g.C32 = dC;
  g.ldc = N;
  MHIP(mlp_gemm(g, nullptr));
  MHIP(hipDeviceSynchronize());
  MHIP(hipMemcpy(C, dC, static_cast<size_t>(M) * N * 4, hipMemcpyDeviceToHost));
  return EPI_OK;
}

}  // extern "C"
