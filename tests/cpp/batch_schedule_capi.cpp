// C entry points over the batch schedule (spla_amd/csrc/runtime/batch_schedule.cpp)
// for the host tests (tests/test_batch_schedule_host.py builds this with g++).
#include <cstring>
#include <vector>

#include "runtime/batch_schedule.hpp"

extern "C" {

// shapes: count rows of (m, n, k, fillMode, rowOff, colOff)
void* bs_plan(const long long* shapes, int count, int complex, int alphaZero, int cfg,
              double splitScale, int numCUs, double efficiency) {
  std::vector<spla_amd::BatchShape> v(count);
  for (int i = 0; i < count; ++i) {
    const long long* s = shapes + 6 * i;
    v[i] = {s[0], s[1], s[2], static_cast<int>(s[3]), s[4], s[5]};
  }
  return new spla_amd::BatchSchedule(spla_amd::plan_batch_schedule(
      v, complex != 0, alphaZero != 0, cfg, splitScale, numCUs, efficiency));
}

// out: items, reduce tiles, workspace doubles, tile configuration
void bs_counts(void* h, long long* out) {
  auto* s = static_cast<spla_amd::BatchSchedule*>(h);
  out[0] = static_cast<long long>(s->items.size());
  out[1] = static_cast<long long>(s->reduce.size());
  out[2] = s->workDoubles;
  out[3] = s->cfg;
}

double bs_model_seconds(void* h) { return static_cast<spla_amd::BatchSchedule*>(h)->modelSeconds; }

// per block: m, n, Kp, Mp, splits, kSplit, workOffset
void bs_blocks(void* h, long long* out) {
  auto* s = static_cast<spla_amd::BatchSchedule*>(h);
  for (std::size_t b = 0; b < s->blocks.size(); ++b) {
    const spla_amd::BatchBlock& x = s->blocks[b];
    long long* o = out + 7 * b;
    o[0] = x.m;
    o[1] = x.n;
    o[2] = x.Kp;
    o[3] = x.Mp;
    o[4] = x.splits;
    o[5] = x.kSplit;
    o[6] = x.workOffset;
  }
}

void bs_items(void* h, int* out) {
  auto* s = static_cast<spla_amd::BatchSchedule*>(h);
  std::memcpy(out, s->items.data(), s->items.size() * sizeof(spla_amd::BatchItem));
}

void bs_reduce(void* h, int* out) {
  auto* s = static_cast<spla_amd::BatchSchedule*>(h);
  std::memcpy(out, s->reduce.data(), s->reduce.size() * sizeof(spla_amd::BatchReduceTile));
}

// static schedule of the persistent launch: slots + 1 starts, then the item list
long long bs_num_slots(void* h) {
  return static_cast<long long>(static_cast<spla_amd::BatchSchedule*>(h)->slotStart.size()) - 1;
}
void bs_slots(void* h, int* start, int* items) {
  auto* s = static_cast<spla_amd::BatchSchedule*>(h);
  std::memcpy(start, s->slotStart.data(), s->slotStart.size() * sizeof(int));
  std::memcpy(items, s->slotItems.data(), s->slotItems.size() * sizeof(int));
}

void bs_free(void* h) { delete static_cast<spla_amd::BatchSchedule*>(h); }
}
