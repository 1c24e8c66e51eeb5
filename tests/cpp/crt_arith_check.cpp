// Runs the FP64 limb CRT of the split-precision reconstruction (crt_arith.hpp,
// the header the kernel includes) on the CPU, with the constants of the
// planner (split_plan.cpp), and compares every result with exact 192-bit
// integer arithmetic:
//   * for every moduli count 1..20, random integers X in the canonical range
//     (-M/2, M/2] of every size, and its edges (0, +-1, M/2, -(M/2 - 1), ...):
//     the reconstruction from the symmetric residues of X must equal the
//     double of X formed from its magnitude's 64-bit words, bit for bit;
//   * the complex combination (Re = T1 - T2, Im = T3 - T1 - T2 modulo m) from
//     biased-byte residues and from int32 split sums;
//   * the symmetric residue of an INT32 sum as the GEMM epilogue forms it
//     (in [-128, 127], congruent to the sum).
// Exit status 0 when all agree.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "kernels/crt_arith.hpp"
#include "runtime/split_plan.hpp"

using spla_amd::CrtConstants;
using spla_amd::ModuliSet;
using spla_amd::crt::I192;

namespace {

std::mt19937_64 rng(11);
long failures = 0;

int bit_length(const I192& a) {
  const uint64_t w[3] = {a.w0, a.w1, a.w2};
  for (int pos = 191; pos >= 0; --pos)
    if ((w[pos / 64] >> (pos % 64)) & 1) return pos + 1;
  return 0;
}

// |x| mod m for a non-negative x
int mod_small(const I192& x, int m) {
  const uint64_t w[3] = {x.w0, x.w1, x.w2};
  unsigned __int128 r = 0;
  for (int i = 2; i >= 0; --i) r = ((r << 64) | w[i]) % static_cast<unsigned>(m);
  return static_cast<int>(r);
}

// symmetric residue of the signed x
int sym_res(const I192& x, int m) {
  const bool neg = spla_amd::crt::negative(x);
  int r = mod_small(neg ? spla_amd::crt::negate(x) : x, m);
  if (neg) r = (m - r) % m;
  if (r > m - 1 - m / 2) r -= m;
  return r;
}

double expected(const I192& x) {  // exact integer -> double from its magnitude's words
  const bool neg = spla_amd::crt::negative(x);
  const I192 a = neg ? spla_amd::crt::negate(x) : x;
  const double d = (std::ldexp(static_cast<double>(a.w2), 128) +
                    std::ldexp(static_cast<double>(a.w1), 64)) +
                   static_cast<double>(a.w0);
  return neg ? -d : d;
}

template <int L>
double reconstruct(const ModuliSet& ms, const CrtConstants& cc, const int* r) {
  double S[L];
  for (int l = 0; l < L; ++l) S[l] = 0.0;
  for (int t = 0; t < ms.count; ++t)
    for (int l = 0; l < L; ++l) S[l] = std::fma(static_cast<double>(r[t]), cc.Wl[t][l], S[l]);
  return spla_amd::crt::finish<L>(S, cc);
}

double reconstruct_any(const ModuliSet& ms, const CrtConstants& cc, const int* r) {
  switch (cc.limbs) {
    case 1: return reconstruct<1>(ms, cc, r);
    case 2: return reconstruct<2>(ms, cc, r);
    case 3: return reconstruct<3>(ms, cc, r);
    default: return reconstruct<4>(ms, cc, r);
  }
}

I192 random_below_bits(int bits) {  // uniform magnitude below 2^bits, random sign
  I192 x{rng(), rng(), rng()};
  uint64_t* w[3] = {&x.w0, &x.w1, &x.w2};
  for (int i = 0; i < 3; ++i) {
    const int lo = 64 * i;
    if (bits <= lo)
      *w[i] = 0;
    else if (bits < lo + 64)
      *w[i] &= (uint64_t(1) << (bits - lo)) - 1;
  }
  return (rng() & 1) ? spla_amd::crt::negate(x) : x;
}

void check_value(const ModuliSet& ms, const CrtConstants& cc, const I192& x) {
  int r[spla_amd::kMaxModuli];
  for (int t = 0; t < ms.count; ++t) r[t] = sym_res(x, ms.m[t]);
  const double got = reconstruct_any(ms, cc, r);
  const double want = expected(x);
  if (!(got == want) || std::signbit(got) != std::signbit(want)) {
    if (failures < 5)
      std::printf("count=%d x=%016llx %016llx %016llx got %.17g want %.17g\n", ms.count,
                  (unsigned long long)x.w2, (unsigned long long)x.w1, (unsigned long long)x.w0,
                  got, want);
    ++failures;
  }
}

// complex combination from biased bytes (in8) or int32 split sums
void check_complex(const ModuliSet& ms, const CrtConstants& cc, int bits, bool in8) {
  const I192 x1 = random_below_bits(bits), x2 = random_below_bits(bits),
             x3 = random_below_bits(bits);
  int re[spla_amd::kMaxModuli], im[spla_amd::kMaxModuli];
  for (int t = 0; t < ms.count; ++t) {
    const int m = ms.m[t];
    double T[3];
    const I192* xs[3] = {&x1, &x2, &x3};
    for (int p = 0; p < 3; ++p) {
      const int sr = sym_res(*xs[p], m);
      if (in8) {
        // the byte the GEMM stores: symmetric residue in [-128, 127] plus 128
        int b = sr;
        if (b > 127) b -= m;
        const uint32_t byte = static_cast<uint32_t>(b + 128) & 0xffu;
        T[p] = static_cast<double>(static_cast<int>(byte) - 128);
      } else {
        // int32 split sums congruent to the residue, split over 3 splits
        const int64_t v = sr + m * static_cast<int64_t>(rng() % 20000001) - m * int64_t(10000000);
        T[p] = static_cast<double>(v);
      }
    }
    const double m_ = m, inv = ms.inv[t];
    re[t] = static_cast<int>(spla_amd::crt::sym_residue(T[0] - T[1], m_, inv));
    im[t] = static_cast<int>(spla_amd::crt::sym_residue(T[2] - T[0] - T[1], m_, inv));
    if (std::abs(re[t]) > m / 2 || std::abs(im[t]) > m / 2) ++failures;
  }
  I192 wr = x1, wi = x3;
  spla_amd::crt::add_in(wr, spla_amd::crt::negate(x2));
  spla_amd::crt::add_in(wi, spla_amd::crt::negate(x1));
  spla_amd::crt::add_in(wi, spla_amd::crt::negate(x2));
  if (!(reconstruct_any(ms, cc, re) == expected(wr)) ||
      !(reconstruct_any(ms, cc, im) == expected(wi))) {
    if (failures < 5) std::printf("complex count=%d in8=%d mismatch\n", ms.count, in8);
    ++failures;
  }
}

void check_epilogue(const ModuliSet& ms) {
  for (int t = 0; t < ms.count; ++t) {
    const int m = ms.m[t];
    const double dm = m, dinv = ms.inv[t], tie = ms.kc[t].tie;
    for (int i = 0; i < 20000; ++i) {
      int64_t v = static_cast<int64_t>(rng() >> 33) - (int64_t(1) << 30);
      if (i < 512) v = (int64_t(1) << 31) - 256 + i;  // largest sums
      if (i >= 512 && i < 1024) v = -(int64_t(1) << 31) + (i - 512);
      if (v > INT32_MAX || v < INT32_MIN) continue;
      const double vd = static_cast<double>(v);
      const double rr = std::fma(-std::rint(std::fma(vd, dinv, tie)), dm, vd);
      const int64_t ri = static_cast<int64_t>(rr);
      if (rr != static_cast<double>(ri) || ri < -128 || ri > 127 || ((v - ri) % m) != 0) {
        if (failures < 5) std::printf("epilogue m=%d v=%lld r=%g\n", m, (long long)v, rr);
        ++failures;
      }
    }
  }
}

}  // namespace

int main() {
  long n = 0;
  for (int count = 1; count <= spla_amd::kMaxModuli; ++count) {
    ModuliSet ms;
    spla_amd::make_moduli(count, ms);
    CrtConstants cc;
    spla_amd::make_crt_constants(ms, cc);
    const I192 M{cc.M[0], cc.M[1], cc.M[2]};
    const I192 half{cc.halfM[0], cc.halfM[1], cc.halfM[2]};
    const I192 low{cc.lowM[0], cc.lowM[1], cc.lowM[2]};
    const int bm = bit_length(M);
    // edges of the canonical range (-M/2, M/2] and around zero
    I192 edges[8] = {{0, 0, 0}, {1, 0, 0}, spla_amd::crt::negate(I192{1, 0, 0}), half,
                     spla_amd::crt::negate(low), half, low, spla_amd::crt::negate(low)};
    spla_amd::crt::add_in(edges[5], spla_amd::crt::negate(I192{1, 0, 0}));  // M/2 - 1
    for (const I192& e : edges) {
      check_value(ms, cc, e);
      ++n;
    }
    for (int i = 0; i < 20000; ++i) {
      const int bits = static_cast<int>(rng() % static_cast<uint64_t>(bm - 1));  // < 2^(bm-2) <= M/2
      check_value(ms, cc, random_below_bits(bits));
      ++n;
    }
    for (int i = 0; i < 4000; ++i) {
      // |x3 - x1 - x2| < 3 2^(bm-4) < 2^(bm-2) <= M/2
      const int bits = std::max(0, bm - 4 - static_cast<int>(rng() % 40));
      check_complex(ms, cc, bits, (i & 1) != 0);
      ++n;
    }
    check_epilogue(ms);
    (void)M;
  }
  std::printf("%ld reconstructions, %ld mismatches\n", n, failures);
  return failures == 0 ? 0 : 1;
}
