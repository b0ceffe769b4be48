// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel Random
// Numbers: As Easy as 1, 2, 3", SC'11), written for both the host (population
// synthesis) and gfx950 device code (the hour step). Ten rounds, each one 32x32->64
// multiply per lane pair; the key is bumped by the paper's Weyl constants between rounds.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define EPI_HD __host__ __device__ __forceinline__
#else
#define EPI_HD inline
#endif

namespace epi {

struct u32x4 {
  uint32_t w[4];
};

// One 32x32 -> 64 product: a single v_mad_u64_u32 on gfx950 (measured 1.32x the Philox
// throughput of a v_mul_lo_u32 + v_mul_hi_u32 pair).
EPI_HD void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = static_cast<uint64_t>(a) * b;
  lo = static_cast<uint32_t>(p);
  hi = static_cast<uint32_t>(p >> 32);
}

// a ^ b ^ c: one v_bitop3_b32 (truth table 0x96) on gfx950, where the compiler otherwise
// emits two v_xor_b32 for each of the round's two three-way xors.
EPI_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

EPI_HD u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c0, hi0, lo0);
    mulhilo32(0xCD9E8D57u, c2, hi1, lo1);
    const uint32_t n0 = xor3(hi1, c1, k0);
    const uint32_t n2 = xor3(hi0, c3, k1);
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return u32x4{{c0, c1, c2, c3}};
}

// One draw block for (seed, id, hour, site, block).
EPI_HD u32x4 draw(uint64_t seed, uint32_t id, uint32_t hour, uint32_t site, uint32_t block) {
  return philox4x32_10(id, hour, site, block, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

EPI_HD uint64_t join64(uint32_t hi, uint32_t lo) { return (static_cast<uint64_t>(hi) << 32) | lo; }

// Uniform integer on [0, n) from one 32-bit word: multiply-shift.
EPI_HD uint32_t below(uint32_t word, uint32_t n) {
  return static_cast<uint32_t>((static_cast<uint64_t>(word) * n) >> 32);
}

// Bernoulli with a precomputed 64-bit threshold (see bernoulli_threshold).
constexpr uint64_t kAlways = ~0ull;
EPI_HD bool bernoulli(uint64_t thr, uint64_t v) { return thr == kAlways || v < thr; }

}  // namespace epi
