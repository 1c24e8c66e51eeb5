#include "runtime/tuner.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kernels/gemm_desc.hpp"

namespace spla_amd {
namespace {

// Per-CU FP64 matrix rate of gfx950: 128 flop / clock / CU at 2.4 GHz.
constexpr double kPeakPerCU = 128.0 * 2.4e9;
constexpr double kHbmBytesPerSecond = 5.0e12;  // sustained, for the split-K reduction
constexpr double kLaunchSeconds = 4.0e-6;

int blocks_per_cu(int cfg) {
  switch (cfg) {
    case TileXL:
      return 1;
    case TileM:
      return 4;
    default:
      return 2;
  }
}

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

int env_tile() {
  const char* v = std::getenv("SPLA_AMD_TILE");
  if (!v) return -1;
  if (!std::strcmp(v, "L")) return TileL;
  if (!std::strcmp(v, "M")) return TileM;
  if (!std::strcmp(v, "XL")) return TileXL;
  return -1;
}

}  // namespace

GemmPlan plan_gemm(bool complex, int64_t m, int64_t n, int64_t k, int numCUs, int forceCfg,
                   int forceSplits, const double* effByCfg, bool wholeChip) {
  const int64_t Mp = complex ? 2 * m : m;
  const int64_t Kp = complex ? 2 * k : k;
  const int cands[] = {TileL, TileM, TileXL};
  GemmPlan best;
  best.estSeconds = 1e300;
  for (int cfg : cands) {
    if (forceCfg >= 0 && cfg != forceCfg) continue;
    if (forceCfg < 0 && cfg == TileXL) continue;  // opt-in only
    const TileShape ts = tile_shape(cfg);
    const int64_t tilesM = (Mp + ts.bm - 1) / ts.bm;
    const int64_t tilesN = (n + ts.bn - 1) / ts.bn;
    const int64_t tiles = tilesM * tilesN;
    const int64_t kBlocks = std::max<int64_t>(1, (Kp + ts.bk - 1) / ts.bk);
    const int bpc = blocks_per_cu(cfg);
    const int64_t slots = static_cast<int64_t>(numCUs) * bpc;
    const double eff = effByCfg ? effByCfg[cfg] : 0.8;
    // split factors: every split keeps at least 8 k-blocks; up to 64, or, for a
    // product that has the device to itself, as many as fill the resident workgroup
    // slots when there are few tiles (small m, n and a long reduction: HBM-bound,
    // every CU must stream its share of A and B). Batch blocks share the device
    // with the blocks on the other streams and keep the cap of 64.
    const int64_t fillSplits = wholeChip ? slots / std::max<int64_t>(1, tiles) : 0;
    const int64_t maxSplits = std::max<int64_t>(
        1, std::min<int64_t>(kBlocks / 8, std::max<int64_t>(64, fillSplits)));
    // A and B are read once whatever the split; with fewer workgroups than resident
    // slots only that share of the HBM bandwidth is drawn
    const double bytes = static_cast<double>(m + n) * static_cast<double>(k) * (complex ? 16.0 : 8.0);
    // candidates 1, 2, 3, 4, 6, 9, ... up to maxSplits and the count that fills every
    // resident slot once, or exactly the forced count
    const int64_t forced = forceSplits > 0 ? std::min<int64_t>(forceSplits, maxSplits) : 0;
    std::vector<int64_t> cand;
    if (forced > 0) {
      cand.push_back(forced);
    } else {
      for (int64_t s = 1; s <= maxSplits; s = (s < 4 ? s + 1 : s * 3 / 2)) cand.push_back(s);
      if (fillSplits > 0 && fillSplits <= maxSplits) cand.push_back(fillSplits);
    }
    for (const int64_t s : cand) {
      const int64_t kbPerSplit = (kBlocks + s - 1) / s;
      const int64_t splits = (kBlocks + kbPerSplit - 1) / kbPerSplit;  // no empty split
      const int64_t blocks = tiles * splits;
      if (blocks > (int64_t(1) << 31) - 1) continue;
      const int64_t waves = (blocks + slots - 1) / slots;
      const double blockFlops = 2.0 * ts.bm * ts.bn * kbPerSplit * ts.bk;
      const double waveSeconds = blockFlops * bpc / (kPeakPerCU * eff);
      // HBM throughput grows with the resident workgroups (their loads in flight)
      const double active = std::min(1.0, static_cast<double>(blocks) / slots);
      const double tMem = wholeChip ? bytes / (kHbmBytesPerSecond * active) : 0.0;
      double t = std::max(waves * waveSeconds, tMem) + kLaunchSeconds;
      if (splits > 1)
        t += (splits + 2) * 8.0 * Mp * n / kHbmBytesPerSecond + kLaunchSeconds;
      if (t < best.estSeconds * 0.999) {
        best.cfg = cfg;
        best.splits = static_cast<int>(splits);
        best.kSplit = kbPerSplit * ts.bk;
        best.tilesM = static_cast<int>(tilesM);
        best.tilesN = static_cast<int>(tilesN);
        best.estSeconds = t;
      }
    }
  }
  if (best.kSplit <= 0) {
    // no candidate passed (e.g. a forced tile whose grid would not fit): one split
    best.cfg = forceCfg >= 0 ? forceCfg : TileL;
    const TileShape ts = tile_shape(best.cfg);
    best.splits = 1;
    best.kSplit = std::max<int64_t>(1, (Kp + ts.bk - 1) / ts.bk) * ts.bk;
    best.tilesM = static_cast<int>((Mp + ts.bm - 1) / ts.bm);
    best.tilesN = static_cast<int>((n + ts.bn - 1) / ts.bn);
  }
  return best;
}

double Tuner::efficiency(int cfg) {
  switch (cfg) {
    case TileXL:
      return 0.85;
    case TileM:
      return 0.55;
    default:
      return 0.8;
  }
}

GemmPlan Tuner::plan(bool complex, int layoutA, int layoutB, int64_t m, int64_t n, int64_t k,
                     int numCUs, bool wholeChip) {
  const int forceCfg = env_tile();
  const int forceSplits = env_int("SPLA_AMD_SPLITS", 0);
  const Key key{complex, layoutA, layoutB, m, n, k, numCUs, wholeChip};
  std::lock_guard<std::mutex> guard(lock_);
  if (forceCfg < 0 && forceSplits <= 0) {
    auto it = cache_.find(key);
    if (it != cache_.end()) return it->second;
  }
  double eff[NumTileConfigs];
  for (int c = 0; c < NumTileConfigs; ++c) eff[c] = efficiency(c);
  GemmPlan p = plan_gemm(complex, m, n, k, numCUs, forceCfg, forceSplits, eff, wholeChip);
  if (forceCfg < 0 && forceSplits <= 0) cache_[key] = p;
  return p;
}

}  // namespace spla_amd
