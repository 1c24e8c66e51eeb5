#include "runtime/split_plan.hpp"

#include "kernels/residue_arith.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace spla_amd {

const int kModuliList[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                     223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

double fp64_bound(int64_t k, int comps) {
  return static_cast<double>(k) * comps * std::ldexp(1.0, -53);
}

namespace {

// ilogb of a positive finite double: the exponent field for normal numbers (the
// library for subnormals). The planner runs between the statistics pass and the
// residue pass, so its per-row work is kept to bit operations.
int ilogb_pos(double x) {
  uint64_t b;
  std::memcpy(&b, &x, sizeof b);
  const int e = static_cast<int>((b >> 52) & 0x7ff);
  return e != 0 ? e - 1023 : std::ilogb(x);
}

// 2^e exactly, built from the bits in the normal range
double pow2(int e) {
  if (e < -1022 || e > 1023) return std::ldexp(1.0, e);
  const uint64_t b = static_cast<uint64_t>(e + 1023) << 52;
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

struct RowScan {
  bool finite = true;
  bool zero = true;
  int maxExp = std::numeric_limits<int>::min();
  double normSq = 0.0;  // ||X||_F^2
};

RowScan scan(const OperandStats& s) {
  RowScan r;
  for (int i = 0; i < s.rows; ++i) {
    const double mx = s.maxabs[i];
    if (!std::isfinite(mx) || !std::isfinite(s.sumsq[i])) {
      r.finite = false;
      continue;
    }
    if (mx > 0.0) {
      r.zero = false;
      r.maxExp = std::max(r.maxExp, ilogb_pos(mx));
    }
    r.normSq += s.sumsq[i];
  }
  return r;
}

// S such that rho(P) = 2^-P * S (see header); computed relative to 2^maxExp.
double scale_factor(const OperandStats& s, const RowScan& r) {
  double acc = 0.0;
  for (int i = 0; i < s.rows; ++i)
    if (s.maxabs[i] > 0.0) acc += pow2(2 * (ilogb_pos(s.maxabs[i]) - r.maxExp));
  const double comps_k = static_cast<double>(s.comps) * static_cast<double>(s.k);
  // sqrt(comps k sum_i 4^fl_i) / ||X|| with 4^fl_i = 4^maxExp * 4^(fl_i - maxExp)
  return std::ldexp(std::sqrt(comps_k * acc), r.maxExp) / std::sqrt(r.normSq);
}

void exponents(const OperandStats& s, int P, std::vector<int>& e) {
  e.assign(s.rows, 0);
  for (int i = 0; i < s.rows; ++i)
    if (s.maxabs[i] > 0.0) e[i] = P - 1 - ilogb_pos(s.maxabs[i]);
}

// log2 of an upper bound of the Euclidean norm of the scaled integer rows:
// max_i min(cap, log2(2^e_i ||x_i|| + 0.5 sqrt(comps k))). The estimate grows with
// 2^(2 e_i) sumsq_i, so only the row where that is largest (compared exactly, as
// 2 e_i + exponent and significand of sumsq_i) is evaluated; a row with a positive
// maximum but a zero sum of squares (underflow) takes the cap, the largest value.
double max_row_norm_log2(const OperandStats& s, const std::vector<int>& e, int P) {
  const double comps_k = static_cast<double>(s.comps) * static_cast<double>(s.k);
  const double capLog2 = P + 0.5 * std::log2(comps_k);  // every |a| <= 2^P
  int best = -1;
  int64_t bestKey = 0;
  double bestSig = 0.0;
  for (int i = 0; i < s.rows; ++i) {
    if (!(s.maxabs[i] > 0.0)) continue;
    const double sq = s.sumsq[i];
    if (!(sq > 0.0)) return capLog2;
    int ex;
    const double sig = std::frexp(sq, &ex);  // sq = sig 2^ex, sig in [0.5, 1)
    const int64_t key = 2 * static_cast<int64_t>(e[i]) + ex;
    if (best < 0 || key > bestKey || (key == bestKey && sig > bestSig)) {
      best = i;
      bestKey = key;
      bestSig = sig;
    }
  }
  if (best < 0) return -std::numeric_limits<double>::infinity();
  // 2^e ||x|| + 0.5 sqrt(comps k), in log2
  const double a = e[best] + 0.5 * std::log2(s.sumsq[best]);
  const double b = std::log2(0.5 * std::sqrt(comps_k));
  const double hi = std::max(a, b), lo = std::min(a, b);
  return std::min(capLog2, hi + std::log2(1.0 + std::exp2(lo - hi)));
}

}  // namespace

SplitPlan plan_split(const OperandStats& A, const OperandStats& B, double tolerance,
                     int maxModuli) {
  SplitPlan p;
  p.tolerance = tolerance;
  const RowScan ra = scan(A), rb = scan(B);
  if (!ra.finite || !rb.finite) {
    p.reason = SplitNonFinite;
    return p;
  }
  if (ra.zero || rb.zero) {
    p.reason = SplitZeroOperand;
    return p;
  }
  if (!(ra.normSq > 0.0) || !std::isfinite(ra.normSq) || !(rb.normSq > 0.0) ||
      !std::isfinite(rb.normSq)) {
    p.reason = SplitRange;
    return p;
  }
  const double SA = scale_factor(A, ra), SB = scale_factor(B, rb);
  if (!std::isfinite(SA) || !std::isfinite(SB)) {
    p.reason = SplitRange;
    return p;
  }
  // rho_A + rho_B + rho_A rho_B <= tol with rho_A, rho_B <= tol / 2.02
  const double half = tolerance / 2.02;
  p.PA = std::max(1, static_cast<int>(std::ceil(std::log2(SA / half))));
  p.PB = std::max(1, static_cast<int>(std::ceil(std::log2(SB / half))));
  p.rhoA = std::ldexp(SA, -p.PA);
  p.rhoB = std::ldexp(SB, -p.PB);
  p.bound = p.rhoA + p.rhoB + p.rhoA * p.rhoB;
  if (p.PA > 53 || p.PB > 53) {
    p.reason = SplitTooWide;
    return p;
  }
  exponents(A, p.PA, p.expA);
  exponents(B, p.PB, p.expB);
  // M > 2 max |a_i . b_j| with margin for the estimate itself
  p.bitsNeeded = max_row_norm_log2(A, p.expA, p.PA) + max_row_norm_log2(B, p.expB, p.PB) + 2.0;
  double bits = 0.0;
  int n = 0;
  while (n < kMaxModuli && bits < p.bitsNeeded) bits += std::log2(kModuliList[n++]);
  if (bits < p.bitsNeeded || n > maxModuli) {
    p.reason = SplitTooManyModuli;
    p.moduli = n;
    return p;
  }
  p.moduli = n;
  p.ok = true;
  p.reason = SplitOk;
  return p;
}

namespace {

int64_t round_up_to(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// HBM rate of the INT32 split sums (written by the GEMM, read by the CRT)
constexpr double kSumBytesPerSecond = 4.0e12;

I8Shape i8_shape(int cfg, int64_t m, int64_t n, int64_t kpad, int G, int numCUs) {
  I8Shape sh;
  sh.cfg = cfg;
  i8_tile_shape(cfg, sh.bm, sh.bn);
  sh.tiles = ((m + sh.bm - 1) / sh.bm) * ((n + sh.bn - 1) / sh.bn);
  // one 256 x 256 workgroup per CU, two 128 x 128 ones
  const int64_t slots = static_cast<int64_t>(numCUs) * (cfg == I8TileL ? 1 : 2);
  int64_t splits = (kpad + kMaxSplitK - 1) / kMaxSplitK;
  while (sh.tiles * G * splits < 2 * numCUs && kpad / (splits + 1) >= 2048) ++splits;
  if (splits > 1) {
    // The GEMM takes ceil(blocks / slots) waves of kpad / splits each. Already
    // summing INT32 splits, up to three more splits are taken when they fill the
    // last wave better (e.g. 15 products x 64 tiles: 2 splits leave the 8th wave
    // half empty, 4 splits make exactly 15 full waves).
    auto cost = [&](int64_t sp) {
      return std::ceil(static_cast<double>(sh.tiles * G * sp) / slots) / static_cast<double>(sp);
    };
    int64_t best = splits;
    for (int64_t sp = splits + 1; sp <= splits + 3 && kpad / sp >= 2048; ++sp)
      if (cost(sp) < 0.97 * cost(best)) best = sp;
    splits = best;
  }
  sh.kSplit = round_up_to((kpad + splits - 1) / splits, kKpadAlign);
  sh.splits = (kpad + sh.kSplit - 1) / sh.kSplit;
  // whole tiles at the shape's rate (128 x 128 tiles read twice the LDS bytes per
  // MFMA: ~0.58 of the 256 x 256 rate), the last wave's fill, ~6.5 k-blocks of
  // prologue and epilogue per workgroup, and the INT32 split sums
  const double blocks = static_cast<double>(sh.tiles) * G * sh.splits;
  const double fill = blocks / (std::ceil(blocks / slots) * slots);
  const double nk = static_cast<double>(sh.kSplit) / 128.0;
  const double ops = 2.0 * G * static_cast<double>(((m + sh.bm - 1) / sh.bm) * sh.bm) *
                     static_cast<double>(((n + sh.bn - 1) / sh.bn) * sh.bn) *
                     static_cast<double>(kpad);
  const double rate = kInt8OpsPerSecond * (cfg == I8TileL ? 1.0 : 0.58);
  const double sums = static_cast<double>(G) * m * n * (sh.splits == 1 ? 1.0 : 8.0 * sh.splits);
  sh.seconds = ops / (rate * fill * nk / (nk + 6.5)) + sums / kSumBytesPerSecond;
  return sh;
}

}  // namespace

I8Shape plan_i8_shape(int64_t m, int64_t n, int64_t kpad, int G, int numCUs) {
  // More splits of 256 x 256 tiles beat 128 x 128 tiles on mid-size products
  // (512 x 512 x 2e5 complex: 1.98 against 2.9 ms); 128 x 128 tiles win where
  // 256-row padding would multiply the work.
  const I8Shape large = i8_shape(I8TileL, m, n, kpad, G, numCUs);
  const I8Shape medium = i8_shape(I8TileM, m, n, kpad, G, numCUs);
  return medium.seconds < large.seconds ? medium : large;
}

int min_moduli(double tolerance) {
  // scale_factor: ||X||^2 < comps k sum_i 4^(fl_i + 1), so S > 1/2 and
  // P = ceil(log2(S / (tol / 2.02))) >= ceil(log2(1.01 / tol)); bitsNeeded is
  // at least (P_A - 1) + (P_B - 1) + 2 (a scaled row's norm is >= 2^(P - 1))
  const int P = std::max(1, static_cast<int>(std::ceil(std::log2(1.01 / tolerance))));
  double bits = 0.0;
  int n = 0;
  while (n < kMaxModuli && bits < 2.0 * P) bits += std::log2(kModuliList[n++]);
  return n;
}

void make_moduli(int count, ModuliSet& ms) {
  std::memset(&ms, 0, sizeof(ms));
  ms.count = count;
  for (int t = 0; t < count; ++t) {
    const int m = kModuliList[t];
    ms.m[t] = m;
    ms.inv[t] = 1.0 / m;
    ms.kc[t] = residue::mod_const(m);
  }
}

namespace {

struct U192 {
  uint64_t w[3] = {0, 0, 0};
};

void mul_small(U192& x, uint32_t s) {
  unsigned __int128 carry = 0;
  for (int i = 0; i < 3; ++i) {
    const unsigned __int128 p = static_cast<unsigned __int128>(x.w[i]) * s + carry;
    x.w[i] = static_cast<uint64_t>(p);
    carry = p >> 64;
  }
}

uint32_t divmod_small(U192& x, uint32_t s) {
  unsigned __int128 rem = 0;
  for (int i = 2; i >= 0; --i) {
    const unsigned __int128 cur = (rem << 64) | x.w[i];
    x.w[i] = static_cast<uint64_t>(cur / s);
    rem = cur % s;
  }
  return static_cast<uint32_t>(rem);
}

uint32_t mod_small(U192 x, uint32_t s) { return divmod_small(x, s); }

int64_t inverse_mod(int64_t a, int64_t m) {
  int64_t t = 0, newt = 1, r = m, newr = a % m;
  while (newr != 0) {
    const int64_t q = r / newr;
    const int64_t tt = t - q * newt;
    t = newt;
    newt = tt;
    const int64_t rr = r - q * newr;
    r = newr;
    newr = rr;
  }
  if (t < 0) t += m;
  return t;
}

double to_double(const U192& x) {
  return std::ldexp(static_cast<double>(x.w[2]), 128) + std::ldexp(static_cast<double>(x.w[1]), 64) +
         static_cast<double>(x.w[0]);
}

}  // namespace

void make_crt_constants(const ModuliSet& ms, CrtConstants& cc) {
  std::memset(&cc, 0, sizeof(cc));
  U192 M;
  M.w[0] = 1;
  for (int t = 0; t < ms.count; ++t) mul_small(M, static_cast<uint32_t>(ms.m[t]));
  for (int t = 0; t < ms.count; ++t) {
    U192 Mt = M;
    divmod_small(Mt, static_cast<uint32_t>(ms.m[t]));  // exact division
    const uint32_t r = mod_small(Mt, static_cast<uint32_t>(ms.m[t]));
    const int64_t y = inverse_mod(r, ms.m[t]);
    mul_small(Mt, static_cast<uint32_t>(y));
    for (int i = 0; i < 3; ++i) cc.W[t][i] = Mt.w[i];
  }
  for (int i = 0; i < 3; ++i) cc.M[i] = M.w[i];
  U192 h = M;
  divmod_small(h, 2);
  for (int i = 0; i < 3; ++i) cc.halfM[i] = h.w[i];
  // lowM = M - 1 - halfM
  unsigned __int128 borrow = 0;
  for (int i = 0; i < 3; ++i) {
    const unsigned __int128 sub = static_cast<unsigned __int128>(h.w[i]) + borrow + (i == 0);
    const unsigned __int128 mw = M.w[i];
    cc.lowM[i] = static_cast<uint64_t>(mw - sub);
    borrow = mw < sub ? 1 : 0;
  }
  cc.Md = to_double(M);
  // 40-bit limbs of W_t and M, and 2^(40 j) / M
  auto limb = [](const U192& v, int j) {
    const int lo = kCrtLimbBits * j;  // bits lo .. lo + 39 of v
    uint64_t bits = 0;
    for (int b = 0; b < kCrtLimbBits; ++b) {
      const int pos = lo + b;
      if (pos < 192 && ((v.w[pos / 64] >> (pos % 64)) & 1)) bits |= uint64_t(1) << b;
    }
    return static_cast<double>(bits);
  };
  int bitsM = 0;
  for (int pos = 191; pos >= 0; --pos)
    if ((M.w[pos / 64] >> (pos % 64)) & 1) {
      bitsM = pos + 1;
      break;
    }
  cc.limbs = std::max(1, (bitsM + kCrtLimbBits - 1) / kCrtLimbBits);
  for (int t = 0; t < ms.count; ++t) {
    U192 w;
    for (int i = 0; i < 3; ++i) w.w[i] = cc.W[t][i];
    for (int j = 0; j < kMaxCrtLimbs; ++j) cc.Wl[t][j] = limb(w, j);
  }
  for (int j = 0; j < kMaxCrtLimbs; ++j) {
    cc.Ml[j] = limb(M, j);
    cc.cq[j] = std::ldexp(1.0, kCrtLimbBits * j) / cc.Md;
  }
}

}  // namespace spla_amd
