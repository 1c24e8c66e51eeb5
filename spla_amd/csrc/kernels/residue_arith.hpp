// Exact residue arithmetic of the split-precision operands, shared by the
// residue kernels (ozaki.hip), the planner (moduli constants) and the host
// check that runs it on the CPU.
//
// An operand entry scaled to an integer a (|a| <= 2^53, held exactly in a
// double) is cut once into symmetric 14-bit pieces held as floats,
//   a = p3 2^42 + p2 2^28 + p1 2^14 + p0,  |p0|, |p1|, |p2| <= 2^13, |p3| <= 2^11.
// Its residue modulo each m in [128, 256] is then formed in FP32 from
//   x = p0 + p1 c14 + p2 c28 + p3 c42   (c_s = symmetric residue of 2^s mod m,
//                                        |c_s| <= 127; all zero for m = 256),
// an integer with |x| <= 2^13 + 2 * 2^13 * 127 + 2^11 * 127 < 2^21.17, so every
// partial sum is exact in float. For odd m, x / m is never closer than 1 / (2m)
// to a half-integer, and x * float(1 / m) misses x / m by at most
// |x| 2^-23 / m < 0.29 / m: q = rint(x * float(1 / m)) is round(x / m) and
// r = x - q m (one exact FMA) is the symmetric residue in [-(m-1)/2, (m-1)/2]
// with no correction step. For m = 256 every step is exact and a small
// offset (2^-10) sends ties up, giving r in [-128, 127].
//
// Biased form: with x' = x + 128 (128 added once to p0) and the constant
// bias = float(-128 / m), q = rint(x' * float(1 / m) + bias) is still round(x / m)
// (error < (2 |x| + 256) 2^-24 / m + ulp < 0.29 / m) and r + 128 = x' - q m lies in
// [0, 255]: the kernels convert it straight to an unsigned byte and flip the
// top bit of every byte, which is the two's-complement int8 of r.
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

struct Pieces {
  float p3, p2, p1, p0;
};

SPLA_AMD_RES_FN Pieces pieces14(double a) {
  const double h3 = std::rint(a * 0x1p-42);
  const double r3 = std::fma(-h3, 0x1p42, a);  // [-2^41, 2^41], exact
  const double h2 = std::rint(r3 * 0x1p-28);
  const double r2 = std::fma(-h2, 0x1p28, r3);  // [-2^27, 2^27], exact
  const double h1 = std::rint(r2 * 0x1p-14);
  return {static_cast<float>(h3), static_cast<float>(h2), static_cast<float>(h1),
          static_cast<float>(std::fma(-h1, 16384.0, r2))};
}

// Per-modulus constants.
struct ModConst {
  float m;
  float invf;  // float(1 / m)
  float tie;   // offset of the unbiased quotient: 2^-10 for m = 256, else 0
  float bias;  // offset of the biased quotient: tie - 128 / m, rounded to float
  float c14, c28, c42;
};

// Symmetric residue of 2^s modulo m.
inline int sym_pow2(int s, int m) {
  int64_t v = (int64_t(1) << s) % m;
  if (v > m - 1 - m / 2) v -= m;
  return static_cast<int>(v);
}

inline ModConst mod_const(int m) {
  const bool even = (m & 1) == 0;  // only m = 256 in the moduli list
  const double tie = even ? 0x1p-10 : 0.0;
  return {static_cast<float>(m),
          1.0f / static_cast<float>(m),
          static_cast<float>(tie),
          static_cast<float>(tie - 128.0 / m),
          static_cast<float>(sym_pow2(14, m)),
          static_cast<float>(sym_pow2(28, m)),
          static_cast<float>(sym_pow2(42, m))};
}

// x (or x' when p0 carries the +128) of the pieces.
SPLA_AMD_RES_FN float mix(const Pieces& s, float p0, const ModConst& k) {
  return std::fma(s.p3, k.c42, std::fma(s.p2, k.c28, std::fma(s.p1, k.c14, p0)));
}

// Symmetric residue r of x (|x| < 2^21.17), as a float.
SPLA_AMD_RES_FN float residue(float x, const ModConst& k) {
  const float q = rintf(std::fma(x, k.invf, k.tie));
  return std::fma(-q, k.m, x);
}

// r + 128 in [0, 255] for x' = x + 128 (|x| < 2^21.17), as a float.
SPLA_AMD_RES_FN float residue_biased(float xb, const ModConst& k) {
  const float q = rintf(std::fma(xb, k.invf, k.bias));
  return std::fma(-q, k.m, xb);
}

// The int8 byte of a biased residue rb in [0, 255].
SPLA_AMD_RES_FN uint32_t biased_byte(float rb) {
  return (static_cast<uint32_t>(rb) ^ 0x80u) & 0xffu;
}

}  // namespace residue
}  // namespace spla_amd
