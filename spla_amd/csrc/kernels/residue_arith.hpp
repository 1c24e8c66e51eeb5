// Exact residue arithmetic of the split-precision operands, shared by the
// residue kernels (ozaki.hip) and the host check that runs it on the CPU.
//
// An operand entry scaled to an integer a (|a| < 2^53, held exactly in a
// double) is cut into 16-bit pieces once; its symmetric residue modulo each
// m in [128, 256] is then formed in 32-bit integer / float arithmetic:
//   x = p0 + p1 c16 + p2 c32 + p3 c48,  c_s = 2^s mod m,  |x| < 2^25
//   q = rint(float(x) * float(1 / m)),  r = x - q m
// float(x) is within 2 of x and float(1 / m) within 2^-24 relative, so
// |x / m - float(x) float(1 / m)| < 0.06 and r lies in (-m, m); one correction
// each way gives the exact residue in [-(m / 2), m - 1 - m / 2].
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
  int p3, p2, p1, p0;  // a = p3 2^48 + p2 2^32 + p1 2^16 + p0, p3 in [-32, 32)
};

SPLA_AMD_RES_FN Pieces pieces16(double a) {
  const double h3 = std::floor(a * 0x1p-48);
  const double r3 = std::fma(-h3, 0x1p48, a);  // [0, 2^48), exact
  const double h2 = std::floor(r3 * 0x1p-32);
  const double r2 = std::fma(-h2, 0x1p32, r3);  // [0, 2^32), exact
  const double h1 = std::floor(r2 * 0x1p-16);
  return {static_cast<int>(h3), static_cast<int>(h2), static_cast<int>(h1),
          static_cast<int>(std::fma(-h1, 65536.0, r2))};
}

// Product of two integers that fit in 24 bits (the result fits in 32).
SPLA_AMD_RES_FN int mul24(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __mul24(a, b);
#else
  return a * b;
#endif
}

SPLA_AMD_RES_FN int sym_mod4(const Pieces& s, int m, int c16, int c32, int c48, float invf, int lo,
                             int hi) {
  const int x = s.p0 + mul24(s.p1, c16) + mul24(s.p2, c32) + mul24(s.p3, c48);
  const int q = static_cast<int>(rintf(static_cast<float>(x) * invf));
  int r = x - mul24(q, m);
  r = (r > hi) ? r - m : r;
  return (r < lo) ? r + m : r;
}

// (r0 + r1) mod m for residues already in [lo, hi].
SPLA_AMD_RES_FN int sym_fold(int r, int m, int lo, int hi) {
  r = (r > hi) ? r - m : r;
  return (r < lo) ? r + m : r;
}

}  // namespace residue
}  // namespace spla_amd
