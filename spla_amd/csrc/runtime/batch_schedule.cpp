#include "runtime/batch_schedule.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <queue>

#include "kernels/gemm_desc.hpp"
#include "spla/types.h"

namespace spla_amd {
namespace {

constexpr double kPeakPerCU = 128.0 * 2.4e9;  // FP64 matrix flop / s per CU
constexpr double kItemOverhead = 2.0e-6;      // prologue + epilogue per workgroup
constexpr double kLaunchOverhead = 8.0e-6;    // per launch in the graph
constexpr double kReduceBytesPerSecond = 4.0e12;
constexpr int kMinKBlocksPerSplit = 16;

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

int batch_blocks_per_cu(int cfg) { return cfg == TileM ? 4 : 2; }

int batch_slots(int cfg, int numCUs, std::size_t items) {
  const std::size_t slots = static_cast<std::size_t>(std::max(1, numCUs)) * batch_blocks_per_cu(cfg);
  return static_cast<int>(std::max<std::size_t>(1, std::min(slots, items)));
}

bool batch_tile_active(int fillMode, int64_t rowOff, int64_t colOff, int64_t r, int64_t c,
                       int64_t nr, int64_t nc) {
  if (fillMode == SPLA_FILL_MODE_UPPER) return c + nc + colOff > r + rowOff;
  if (fillMode == SPLA_FILL_MODE_LOWER) return r + nr + rowOff > c + colOff;
  return true;
}

BatchSchedule plan_batch_schedule(const std::vector<BatchShape>& shapes, bool complex,
                                  bool alphaZero, int cfg, double splitScale, int numCUs,
                                  double efficiency) {
  BatchSchedule out;
  out.cfg = cfg;
  out.splitScale = splitScale;
  const TileShape ts = tile_shape(cfg);
  const int comps = complex ? 2 : 1;
  const int ar = ts.bm / comps;  // element rows per tile
  const int perCU = batch_blocks_per_cu(cfg);
  const int slots = std::max(1, numCUs) * perCU;
  const double slotRate = kPeakPerCU * efficiency / perCU;

  struct Tile {
    int block, tm, tn;
  };
  std::vector<Tile> tiles;
  std::vector<double> tileCost(shapes.size(), 0.0);
  double total = 0.0;
  for (std::size_t b = 0; b < shapes.size(); ++b) {
    const BatchShape& s = shapes[b];
    if (s.m <= 0 || s.n <= 0) continue;
    const int64_t Kp = alphaZero ? 0 : s.k * comps;
    tileCost[b] = 2.0 * ts.bm * ts.bn * static_cast<double>(std::max<int64_t>(Kp, ts.bk)) /
                  slotRate;
    const int64_t tilesM = ceil_div(s.m, ar), tilesN = ceil_div(s.n, ts.bn);
    for (int64_t tn = 0; tn < tilesN; ++tn)
      for (int64_t tm = 0; tm < tilesM; ++tm) {
        const int64_t r = tm * ar, c = tn * ts.bn;
        if (!batch_tile_active(s.fillMode, s.rowOff, s.colOff, r, c, std::min<int64_t>(ar, s.m - r),
                               std::min<int64_t>(ts.bn, s.n - c)))
          continue;
        tiles.push_back({static_cast<int>(b), static_cast<int>(tm), static_cast<int>(tn)});
        total += tileCost[b];
      }
  }
  const double target = std::max(total / (3.0 * slots), 20e-6);

  out.blocks.resize(shapes.size());
  for (std::size_t b = 0; b < shapes.size(); ++b) {
    const BatchShape& s = shapes[b];
    BatchBlock& bb = out.blocks[b];
    std::memset(&bb, 0, sizeof(bb));
    if (s.m <= 0 || s.n <= 0) continue;
    const int64_t Kp = alphaZero ? 0 : s.k * comps;
    const int64_t kBlocks = ceil_div(Kp, ts.bk);
    int64_t sp = static_cast<int64_t>(std::ceil(tileCost[b] * splitScale / target));
    sp = std::max<int64_t>(1, std::min<int64_t>(sp, std::max<int64_t>(1, kBlocks /
                                                                       kMinKBlocksPerSplit)));
    const int64_t kbPer = kBlocks > 0 ? ceil_div(kBlocks, sp) : 1;
    sp = kBlocks > 0 ? ceil_div(kBlocks, kbPer) : 1;  // no empty split
    bb.m = static_cast<int>(s.m);
    bb.n = static_cast<int>(s.n);
    bb.Kp = Kp;
    bb.Mp = static_cast<int>(s.m * comps);
    bb.splits = static_cast<int>(sp);
    bb.kSplit = kbPer * ts.bk;
    if (sp > 1) {
      bb.workOffset = out.workDoubles;
      out.workDoubles += sp * s.n * bb.Mp;
    }
  }

  struct Costed {
    BatchItem item;
    double cost;
  };
  std::vector<Costed> items;
  double reduceSeconds = 0.0;
  for (const Tile& t : tiles) {
    const BatchBlock& bb = out.blocks[t.block];
    for (int q = 0; q < bb.splits; ++q)
      items.push_back({{t.block, t.tm, t.tn, q}, tileCost[t.block] / bb.splits + kItemOverhead});
    if (bb.splits > 1) {
      out.reduce.push_back({t.block, t.tm, t.tn});
      reduceSeconds +=
          (bb.splits + 2.0) * ar * ts.bn * 8.0 * comps / kReduceBytesPerSecond;
    }
  }
  std::stable_sort(items.begin(), items.end(),
                   [](const Costed& a, const Costed& b) { return a.cost > b.cost; });
  out.items.reserve(items.size());
  // longest processing time first: each item goes to the slot that frees up
  // first (ties: the lowest slot), which is also the persistent launch's schedule
  const int nSlots = batch_slots(cfg, numCUs, items.size());
  using Slot = std::pair<double, int>;  // (load, slot)
  std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> load;
  for (int i = 0; i < nSlots; ++i) load.push({0.0, i});
  std::vector<std::vector<int>> perSlot(nSlots);
  double makespan = 0.0;
  for (std::size_t i = 0; i < items.size(); ++i) {
    out.items.push_back(items[i].item);
    Slot sl = load.top();
    load.pop();
    sl.first += items[i].cost;
    makespan = std::max(makespan, sl.first);
    perSlot[sl.second].push_back(static_cast<int>(i));
    load.push(sl);
  }
  out.slotStart.assign(1, 0);
  for (const std::vector<int>& v : perSlot) {
    out.slotItems.insert(out.slotItems.end(), v.begin(), v.end());
    out.slotStart.push_back(static_cast<int>(out.slotItems.size()));
  }
  out.modelSeconds = makespan + reduceSeconds + kLaunchOverhead * (out.reduce.empty() ? 1 : 2);
  return out;
}

}  // namespace spla_amd
