// Division of 32-bit unsigned values by a divisor fixed at set-up, as a multiply-high and
// shifts (Granlund & Montgomery, "Division by invariant integers using multiplication",
// PLDI 1994, the round-up variant): for d >= 2 with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, and q = (t + ((n - t) >> 1)) >> (l - 1), t = mulhi(n, m).
// The step kernel uses it for the house / office tiling (agent id -> building).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define EPI_FD_HD __host__ __device__ __forceinline__
#else
#define EPI_FD_HD inline
#endif

namespace epi {

struct FastDiv {
  uint32_t d;      // divisor, >= 1
  uint32_t m;      // magic multiplier (unused for d == 1)
  uint32_t shift;  // l - 1
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f{d, 0u, 0u};
  if (d <= 1) {
    f.d = 1;
    return f;
  }
  uint32_t l = 0;
  while ((static_cast<uint64_t>(1) << l) < d) ++l;
  f.m = static_cast<uint32_t>(((static_cast<uint64_t>(1) << 32) * ((static_cast<uint64_t>(1) << l) - d)) / d + 1u);
  f.shift = l - 1u;
  return f;
}

EPI_FD_HD uint32_t fd_mulhi(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
#endif
}

EPI_FD_HD uint32_t fast_div(uint32_t n, const FastDiv& f) {
  if (f.d == 1) return n;
  const uint32_t t = fd_mulhi(n, f.m);
  return (t + ((n - t) >> 1)) >> f.shift;
}

EPI_FD_HD uint32_t fast_mod(uint32_t n, const FastDiv& f) { return n - fast_div(n, f) * f.d; }

}  // namespace epi
