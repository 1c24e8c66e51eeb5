// Exact residue arithmetic of the split-precision operands, shared by the
// residue kernels (ozaki.hip) and the host check that runs it on the CPU.
//
// An operand entry scaled to an integer a (|a| <= 2^53, held exactly in a
// double) is cut once into 14-bit pieces held as floats,
//   a = p3 2^42 + p2 2^28 + p1 2^14 + p0,  p0, p1, p2 in [0, 2^14), |p3| <= 2^11.
// Its symmetric residue modulo each m in [128, 256] is then formed in FP32:
//   x = p0 + p1 c14 + p2 c28 + p3 c42   (c_s = symmetric residue of 2^s mod m)
// is an integer with |x| < 2^22.1, so every partial sum is exact in float;
// q = rint(x / m - delta) with delta = 1/200 exceeds the rounding error of
// x * float(1 / m) (< 0.0025), so q is round(x / m) or one less near a tie, and
// r = x - q m lies in [lo, hi + m); one correction gives the residue in
// [lo, hi] = [-(m / 2), m - 1 - m / 2].
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define SPLA_AMD_RES_FN __host__ __device__ inline
#else
#define SPLA_AMD_RES_FN inline
#endif

namespace spla_amd {
namespace residue {

constexpr float kRoundBias = 0.005f;

struct Pieces {
  float p3, p2, p1, p0;
};

SPLA_AMD_RES_FN Pieces pieces14(double a) {
  const double h3 = std::floor(a * 0x1p-42);
  const double r3 = std::fma(-h3, 0x1p42, a);  // [0, 2^42), exact
  const double h2 = std::floor(r3 * 0x1p-28);
  const double r2 = std::fma(-h2, 0x1p28, r3);  // [0, 2^28), exact
  const double h1 = std::floor(r2 * 0x1p-14);
  return {static_cast<float>(h3), static_cast<float>(h2), static_cast<float>(h1),
          static_cast<float>(std::fma(-h1, 16384.0, r2))};
}

// Per-modulus constants of sym_mod.
struct ModConst {
  float m, invf, c14, c28, c42;
  int hi;  // m - 1 - m / 2
};

SPLA_AMD_RES_FN int sym_mod(const Pieces& s, const ModConst& k) {
  const float x = std::fma(s.p3, k.c42, std::fma(s.p2, k.c28, std::fma(s.p1, k.c14, s.p0)));
  const float q = rintf(std::fma(x, k.invf, -kRoundBias));
  const int r = static_cast<int>(std::fma(-q, k.m, x));
  return r > k.hi ? r - static_cast<int>(k.m) : r;
}

// (r0 + r1) mod m for residues already in [lo, hi].
SPLA_AMD_RES_FN int sym_fold(int r, int m, int lo, int hi) {
  r = (r > hi) ? r - m : r;
  return (r < lo) ? r + m : r;
}

// Symmetric residue of 2^s modulo m.
inline int sym_pow2(int s, int m) {
  int64_t v = (int64_t(1) << s) % m;
  if (v > m - 1 - m / 2) v -= m;
  return static_cast<int>(v);
}

}  // namespace residue
}  // namespace spla_amd
