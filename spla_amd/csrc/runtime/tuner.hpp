// Tile-shape and split-K planner of the FP64 MFMA GEMM.
//
// A plan is picked by a cost model of the launch: for every tile shape and
// split factor the model counts waves of workgroups over the CUs, the time of
// one wave at the shape's sustained MFMA rate, tile padding waste, and the
// HBM time of the split-K reduction; the cheapest plan wins. Plans are cached
// per problem key and depend only on the shape, so that repeated calls of one
// product split the reduction alike and return bitwise identical results (the
// batch API's plans, by contrast, can be timed: see batched.hpp).
// Environment overrides (for tuning sweeps): SPLA_AMD_TILE=L|M|XL,
// SPLA_AMD_SPLITS=<n>.
#pragma once

#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>

namespace spla_amd {

struct GemmPlan {
  int cfg = 0;          // GemmTileConfig
  int splits = 1;       // split-K factor
  int64_t kSplit = 0;   // embedded reduction length per split
  int tilesM = 0, tilesN = 0;
  double estSeconds = 0.0;
};

class Tuner {
public:
  // complex: complex operands; m, n, k in elements; numCUs of the device.
  // wholeChip: the product has the device to itself (a single call); false for the
  // block products of a batch, which run side by side on several streams.
  GemmPlan plan(bool complex, int layoutA, int layoutB, int64_t m, int64_t n, int64_t k,
                int numCUs, bool wholeChip = true);

  // Sustained fraction of the per-CU FP64 matrix rate assumed for a tile config.
  static double efficiency(int cfg);

private:
  using Key = std::tuple<bool, int, int, int64_t, int64_t, int64_t, int, bool>;
  std::mutex lock_;
  std::map<Key, GemmPlan> cache_;
};

// Pure cost model (no state), exposed for tests.
GemmPlan plan_gemm(bool complex, int64_t m, int64_t n, int64_t k, int numCUs, int forceCfg,
                   int forceSplits, const double* effByCfg, bool wholeChip = true);

}  // namespace spla_amd
