// C entry point over the FP64 kernels' tile planner, for the host tests
// (tests/test_tuner_host.py builds this with g++ together with
// spla_amd/csrc/runtime/tuner.cpp).
#include "runtime/tuner.hpp"

extern "C" {

// out: [cfg, splits, kSplit, tilesM, tilesN]
void plan_gemm_c(int complex, long long m, long long n, long long k, int numCUs, int forceCfg,
                 int forceSplits, int wholeChip, long long* out) {
  const double eff[8] = {0.8, 0.55, 0.85, 0.8, 0.8, 0.8, 0.8, 0.8};
  spla_amd::GemmPlan p = spla_amd::plan_gemm(complex != 0, m, n, k, numCUs, forceCfg, forceSplits,
                                             eff, wholeChip != 0);
  out[0] = p.cfg;
  out[1] = p.splits;
  out[2] = p.kSplit;
  out[3] = p.tilesM;
  out[4] = p.tilesN;
}
}
