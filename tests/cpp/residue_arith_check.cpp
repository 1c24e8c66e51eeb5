// Runs the residue arithmetic of the split-precision kernels on the CPU and
// compares every residue with exact integer arithmetic: edge values around
// every power of two up to 2^53 (inclusive) and random integers, for every modulus.
// Exit status 0 when all agree.
#include <cstdint>
#include <cstdio>
#include <random>

#include "kernels/residue_arith.hpp"

static const int kModuli[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

static long check(int64_t a) {
  using namespace spla_amd::residue;
  const Pieces p = pieces14(static_cast<double>(a));
  long bad = 0;
  for (int m : kModuli) {
    const int hi = m - 1 - m / 2;
    const ModConst k{static_cast<float>(m), 1.0f / static_cast<float>(m),
                     static_cast<float>(sym_pow2(14, m)), static_cast<float>(sym_pow2(28, m)),
                     static_cast<float>(sym_pow2(42, m)), hi};
    const int got = sym_mod(p, k);
    int64_t r = a % m;
    if (r < 0) r += m;
    if (r > hi) r -= m;
    if (got != r) {
      if (bad < 5) std::printf("a=%lld m=%d got=%d want=%lld\n", (long long)a, m, got, (long long)r);
      ++bad;
    }
  }
  return bad;
}

int main() {
  long bad = 0, n = 0;
  const int64_t lim = int64_t(1) << 53;
  for (int e = 0; e <= 53; ++e)
    for (int64_t d = -3; d <= 3; ++d)
      for (int s = -1; s <= 1; s += 2) {
        const int64_t a = s * ((int64_t(1) << e) + d);
        if (a > lim || a < -lim) continue;
        bad += check(a);
        ++n;
      }
  std::mt19937_64 rng(7);
  for (int i = 0; i < 2000000; ++i) {
    const int bits = 1 + static_cast<int>(rng() % 54);
    int64_t a = static_cast<int64_t>(rng() >> (64 - bits));
    if (a > lim) a = lim;
    if (rng() & 1) a = -a;
    bad += check(a);
    ++n;
  }
  std::printf("%ld values x 20 moduli, %ld mismatches\n", n, bad);
  return bad == 0 ? 0 : 1;
}
