// Runs the residue arithmetic of the split-precision kernels on the CPU and
// compares every residue with exact integer arithmetic: edge values around
// every power of two up to 2^53 (inclusive) and random integers, for every
// modulus. Per value it checks the three planes the complex kernels write:
// the biased residue of a (as its int8 byte), the unbiased residue of a
// partner value b, and the biased residue of their sum, exactly as the
// kernels form them. Exit status 0 when all agree.
#include <cstdint>
#include <cstdio>
#include <random>

#include "kernels/residue_arith.hpp"

static const int kModuli[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

static int64_t sym(int64_t a, int m) {
  int64_t r = a % m;
  if (r < 0) r += m;
  if (r > m - 1 - m / 2) r -= m;
  return r;
}

static long check(int64_t a, int64_t b) {
  using namespace spla_amd::residue;
  Pieces pa = pieces14(static_cast<double>(a));
  const Pieces pb = pieces14(static_cast<double>(b));
  pa.p0 += 128.0f;
  long bad = 0;
  for (int m : kModuli) {
    const ModConst k = mod_const(m);
    const float ra = residue_biased(mix(pa, pa.p0, k), k);
    const float rb = residue(mix(pb, pb.p0, k), k);
    const float rs = residue_biased(ra + rb, k);
    const int ga = static_cast<int8_t>(biased_byte(ra));
    const int gb = static_cast<int8_t>(biased_byte(rb + 128.0f));
    const int gs = static_cast<int8_t>(biased_byte(rs));
    const bool inRange = ra >= 0.f && ra <= 255.f && rb >= -128.f && rb <= 127.f &&
                         rs >= 0.f && rs <= 255.f;
    if (!inRange || ga != sym(a, m) || gb != sym(b, m) || gs != sym(sym(a, m) + sym(b, m), m) ||
        static_cast<float>(ga) != ra - 128.0f || static_cast<float>(gb) != rb) {
      if (bad < 5)
        std::printf("a=%lld b=%lld m=%d got %d %d %d want %lld %lld\n", (long long)a,
                    (long long)b, m, ga, gb, gs, (long long)sym(a, m), (long long)sym(b, m));
      ++bad;
    }
  }
  return bad;
}

int main() {
  long bad = 0, n = 0;
  const int64_t lim = int64_t(1) << 53;
  std::mt19937_64 rng(7);
  auto rnd = [&]() {
    const int bits = 1 + static_cast<int>(rng() % 54);
    int64_t a = static_cast<int64_t>(rng() >> (64 - bits));
    if (a > lim) a = lim;
    return (rng() & 1) ? -a : a;
  };
  for (int e = 0; e <= 53; ++e)
    for (int64_t d = -3; d <= 3; ++d)
      for (int s = -1; s <= 1; s += 2) {
        const int64_t a = s * ((int64_t(1) << e) + d);
        if (a > lim || a < -lim) continue;
        bad += check(a, a) + check(a, -a) + check(a, rnd()) + check(rnd(), a);
        n += 4;
      }
  for (int i = 0; i < 2000000; ++i) {
    bad += check(rnd(), rnd());
    ++n;
  }
  std::printf("%ld value pairs x 20 moduli, %ld mismatches\n", n, bad);
  return bad == 0 ? 0 : 1;
}
