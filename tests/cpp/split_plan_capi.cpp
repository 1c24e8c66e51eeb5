// C entry points over the split-precision planner, for the host tests
// (tests/test_split_plan.py builds this with g++ together with
// spla_amd/csrc/runtime/split_plan.cpp).
#include <cstring>

#include "runtime/split_plan.hpp"

extern "C" {

// out: [ok, reason, PA, PB, moduli]; bound: estimated normwise bound;
// expA / expB: per-row exponents (only written when ok).
void plan_c(const double* maxA, const double* sumA, int rowsA, const double* maxB,
            const double* sumB, int rowsB, long long k, int comps, double tol, int maxModuli,
            int* out, double* bound, int* expA, int* expB) {
  spla_amd::OperandStats a{maxA, sumA, rowsA, k, comps};
  spla_amd::OperandStats b{maxB, sumB, rowsB, k, comps};
  spla_amd::SplitPlan p = spla_amd::plan_split(a, b, tol, maxModuli);
  out[0] = p.ok;
  out[1] = p.reason;
  out[2] = p.PA;
  out[3] = p.PB;
  out[4] = p.moduli;
  *bound = p.bound;
  if (p.ok) {
    std::memcpy(expA, p.expA.data(), rowsA * sizeof(int));
    std::memcpy(expB, p.expB.data(), rowsB * sizeof(int));
  }
}

int min_moduli_c(double tol) { return spla_amd::min_moduli(tol); }

// out: [cfg, bm, bn, tiles, splits, kSplit]; returns the modelled seconds
double i8_shape_c(long long m, long long n, long long kpad, int G, int numCUs, long long* out) {
  const spla_amd::I8Shape s = spla_amd::plan_i8_shape(m, n, kpad, G, numCUs);
  out[0] = s.cfg;
  out[1] = s.bm;
  out[2] = s.bn;
  out[3] = s.tiles;
  out[4] = s.splits;
  out[5] = s.kSplit;
  return s.seconds;
}

// W: count x 3 limbs, M: 3 limbs, halfM: 3 limbs
void crt_c(int count, unsigned long long* W, unsigned long long* M, unsigned long long* halfM,
           double* Md) {
  spla_amd::ModuliSet ms;
  spla_amd::make_moduli(count, ms);
  spla_amd::CrtConstants cc;
  spla_amd::make_crt_constants(ms, cc);
  for (int t = 0; t < count; ++t)
    for (int i = 0; i < 3; ++i) W[3 * t + i] = cc.W[t][i];
  for (int i = 0; i < 3; ++i) {
    M[i] = cc.M[i];
    halfM[i] = cc.halfM[i];
  }
  *Md = cc.Md;
}
}
