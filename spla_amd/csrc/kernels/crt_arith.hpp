// Chinese remainder reconstruction of the split-precision products in FP64
// arithmetic, shared by the reconstruction kernel (ozaki.hip), the planner
// (limb constants) and the host check that runs it on the CPU.
//
// Every W_t (and M) is cut into 40-bit limbs held as doubles,
//   W_t = sum_j W_tj 2^(40 j),  0 <= W_tj < 2^40,  j < L = ceil(bits(M) / 40) <= 4.
// With symmetric residues |r_t| <= 128 of N <= 20 moduli, every limb sum
//   S_j = sum_t r_t W_tj,  |S_j| <= 20 * 2^7 * 2^40 < 2^51.4,
// is an exact FMA chain, and X' = sum_j S_j 2^(40 j) = sum_t r_t W_t is exact
// as the L doubles. Then q = rint(sum_j S_j * (2^(40 j) / M)) (|q| <= 2560,
// estimate error < 2^-36) and D_j = S_j - q M_j (exact: |D_j| < 2^52.4) give
// X' - q M = sum_j D_j 2^(40 j), which is assembled exactly as a 192-bit two's
// complement integer. A last comparison moves it into the canonical range
// (-M/2, M/2] (exactly what the estimate of q can miss, by at most one M), and
// the magnitude is converted to double from its 64-bit words as
// (w2 2^128 + w1 2^64) + w0.
#pragma once

#include <cmath>
#include <cstdint>

#include "kernels/ozaki_desc.hpp"

#ifdef __HIPCC__
#define SPLA_AMD_CRT_FN __host__ __device__ inline
#else
#define SPLA_AMD_CRT_FN inline
#endif

namespace spla_amd {
namespace crt {

// Two's complement 192-bit integer.
struct I192 {
  uint64_t w0, w1, w2;
};

SPLA_AMD_CRT_FN void add_in(I192& a, const I192& b) {
  const uint64_t s0 = a.w0 + b.w0;
  const uint64_t c0 = s0 < a.w0;
  const uint64_t t1 = a.w1 + b.w1;
  const uint64_t c1 = t1 < a.w1;
  const uint64_t s1 = t1 + c0;
  const uint64_t c2 = c1 | (s1 < t1);
  a.w0 = s0;
  a.w1 = s1;
  a.w2 = a.w2 + b.w2 + c2;
}

SPLA_AMD_CRT_FN I192 negate(const I192& a) {
  I192 r{~a.w0, ~a.w1, ~a.w2};
  add_in(r, I192{1, 0, 0});
  return r;
}

SPLA_AMD_CRT_FN bool negative(const I192& a) { return static_cast<int64_t>(a.w2) < 0; }

// d * 2^(40 j) for |d| < 2^63, j < 4.
SPLA_AMD_CRT_FN I192 limb_term(int64_t d, int j) {
  const uint64_t u = static_cast<uint64_t>(d);
  const uint64_t ext = d < 0 ? ~uint64_t(0) : 0;
  switch (j) {
    case 0: return {u, ext, ext};
    case 1: return {u << 40, static_cast<uint64_t>(d >> 24), ext};
    case 2: return {0, u << 16, static_cast<uint64_t>(d >> 48)};
    default: return {0, u << 56, static_cast<uint64_t>(d >> 8)};
  }
}

// Signed value of x as a double: its magnitude's 64-bit words, (w2 2^128 + w1 2^64) + w0.
SPLA_AMD_CRT_FN double to_double(const I192& x) {
  const bool neg = negative(x);
  const I192 a = neg ? negate(x) : x;
  const double d = (std::ldexp(static_cast<double>(a.w2), 128) +
                    std::ldexp(static_cast<double>(a.w1), 64)) +
                   static_cast<double>(a.w0);
  return neg ? -d : d;
}

// Symmetric residue (|r| <= m / 2) of an integer held exactly in a double with
// |x| < 2^45, from q = rint(x / m): x * fl(1 / m) misses x / m by < 2^-7 / m.
SPLA_AMD_CRT_FN double sym_residue(double x, double m, double inv) {
  return std::fma(-std::rint(x * inv), m, x);
}

// The canonical representative in (-M/2, M/2] of sum_j S_j 2^(40 j) mod M,
// converted to double. S holds the L limb sums.
template <int L>
SPLA_AMD_CRT_FN double finish(const double (&S)[L], const CrtConstants& cc) {
  double e = 0.0;
  for (int j = L - 1; j >= 0; --j) e = std::fma(S[j], cc.cq[j], e);
  const double q = std::rint(e);
  I192 x{0, 0, 0};
  for (int j = 0; j < L; ++j)
    add_in(x, limb_term(static_cast<int64_t>(std::fma(-q, cc.Ml[j], S[j])), j));
  const I192 M{cc.M[0], cc.M[1], cc.M[2]};
  I192 hi = negate(x);  // halfM - x
  add_in(hi, I192{cc.halfM[0], cc.halfM[1], cc.halfM[2]});
  if (negative(hi)) {
    add_in(x, negate(M));  // x > M/2
  } else {
    I192 lo = x;  // x + (M - 1 - halfM) < 0: x <= -M/2
    add_in(lo, I192{cc.lowM[0], cc.lowM[1], cc.lowM[2]});
    if (negative(lo)) add_in(x, M);
  }
  return to_double(x);
}

}  // namespace crt
}  // namespace spla_amd
