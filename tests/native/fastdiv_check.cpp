// Checks fast_div / fast_mod (epirust_amd/csrc/fastdiv.hpp) against the hardware division:
// every divisor 1..4096 with numerators 0..70000 and near 2^32, then random divisors up to
// 2^32 - 1 with random and edge numerators. Prints "ok" or the first mismatch.
#include <cstdint>
#include <cstdio>
#include <random>

#include "fastdiv.hpp"

using epi::FastDiv;

static bool check(uint32_t n, const FastDiv& f) {
  if (epi::fast_div(n, f) != n / f.d || epi::fast_mod(n, f) != n % f.d) {
    std::printf("mismatch n=%u d=%u: %u %u\n", n, f.d, epi::fast_div(n, f), epi::fast_mod(n, f));
    return false;
  }
  return true;
}

int main() {
  for (uint32_t d = 1; d <= 4096; ++d) {
    const FastDiv f = epi::make_fastdiv(d);
    for (uint32_t n = 0; n <= 70000; ++n)
      if (!check(n, f)) return 1;
    for (uint32_t k = 0; k < 70000; ++k)
      if (!check(0xFFFFFFFFu - k, f)) return 1;
  }
  std::mt19937_64 rng(7);
  for (int i = 0; i < 200000; ++i) {
    uint32_t d = static_cast<uint32_t>(rng()) >> (rng() % 32);
    if (d == 0) d = 1;
    const FastDiv f = epi::make_fastdiv(d);
    const uint32_t edges[6] = {0u, d - 1u, d, d + 1u, 0xFFFFFFFFu, 0xFFFFFFFFu - d};
    for (uint32_t n : edges)
      if (!check(n, f)) return 1;
    for (int j = 0; j < 50; ++j)
      if (!check(static_cast<uint32_t>(rng()), f)) return 1;
  }
  std::printf("ok\n");
  return 0;
}
