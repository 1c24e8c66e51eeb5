#include "runtime/batched.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <queue>
#include <type_traits>

#include "kernels/gemm_desc.hpp"
#include "kernels/launch.hpp"
#include "runtime/engine.hpp"
#include "runtime/hip_check.hpp"
#include "runtime/split_gemm.hpp"
#include "spla/exceptions.hpp"

namespace spla_amd {
namespace {

constexpr double kPeakPerCU = 128.0 * 2.4e9;  // FP64 matrix flop / s per CU
constexpr double kItemOverhead = 2.0e-6;      // scheduling + epilogue per item
constexpr double kLaunchOverhead = 8.0e-6;

// The persistent launch of batched.hip (one hipGraph node per batch) is
// switched off: its first run on an MI355X produced no result within three
// minutes and the cause has not been identified. Until it is, execute() runs
// the blocks one by one on the single-product engine (same kernels as gemm).
constexpr bool kPersistentLaunch = false;
constexpr double kWorkgroupBytesPerSecond = 60e9;  // one workgroup reducing split partials

int blocks_per_cu(int cfg) { return cfg == TileM ? 4 : 2; }
double efficiency(int cfg) { return cfg == TileM ? 0.55 : 0.75; }

bool tile_active(int mode, int64_t rowOff, int64_t colOff, int64_t r, int64_t c, int64_t nr,
                 int64_t nc) {
  // SPLA block_is_active (src/block_generation/block.hpp:58-66).
  if (mode == SPLA_FILL_MODE_UPPER) return c + nc + colOff > r + rowOff;
  if (mode == SPLA_FILL_MODE_LOWER) return r + nr + rowOff > c + colOff;
  return true;
}

class DeviceGuard {
public:
  explicit DeviceGuard(int dev) {
    hip_check(hipGetDevice(&prev_));
    if (prev_ != dev) {
      hip_check(hipSetDevice(dev));
      changed_ = true;
    }
  }
  ~DeviceGuard() {
    if (changed_) (void)hipSetDevice(prev_);
  }

private:
  int prev_ = 0;
  bool changed_ = false;
};

bool on_device(const void* p) {
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice;
}

// Split-precision batch in three rounds over the lanes: the statistics of
// every block, one synchronization and the host plans of all blocks, then the
// residue planes, INT8 GEMMs and reconstructions (blocks that fall back run
// the FP64 kernels). Each lane reuses one workspace sized for its largest
// block, so the planes of at most one block per lane are held at a time.
template <typename T>
void execute_split_batch(spla::ContextInternal& ctx, const std::vector<BlockSpec>& specs,
                         SplaOperation opA, T alpha, T beta, const hipStream_t* lanes,
                         int nLanes, std::vector<Block>& keep) {
  struct Item {
    const BlockSpec* spec;
    std::unique_ptr<SplitProduct<T>> p;
    int lane;
    bool split;
  };
  const bool alphaZero = alpha == T(0);
  std::vector<Item> items;
  int next = 0;
  for (const BlockSpec& b : specs) {
    if (b.m == 0 || b.n == 0) continue;
    Item it;
    it.spec = &b;
    it.lane = next;
    next = (next + 1) % nLanes;
    FillSpec fill;
    fill.mode = b.fillMode;
    fill.rowOff = b.rowOffset;
    fill.colOff = b.colOffset;
    it.p = std::make_unique<SplitProduct<T>>(
        opA, SPLA_OP_NONE, b.m, b.n, b.k, alpha, static_cast<const T*>(b.A), b.lda,
        static_cast<const T*>(b.B), b.ldb, beta, static_cast<T*>(b.C), b.ldc, ctx, fill);
    // alpha == 0 is beta scaling only, which the FP64 path does without reading A, B;
    // in automatic mode blocks the split form cannot win skip the statistics
    const bool auto_ = ctx.gemm_mode() == GemmModeAuto;
    it.split = !alphaZero &&
               (!auto_ || it.p->may_beat_fp64(fp64_estimate_seconds(
                              ctx, !std::is_same<T, double>::value, opA, SPLA_OP_NONE, b.m, b.n,
                              b.k, /*wholeChip=*/false))) &&
               it.p->enqueue_stats(lanes[it.lane]);
    items.push_back(std::move(it));
  }
  for (int i = 0; i < nLanes; ++i) hip_check(hipStreamSynchronize(lanes[i]));
  std::vector<std::size_t> laneBytes(nLanes, 0);
  const bool cplx = !std::is_same<T, double>::value;
  for (Item& it : items) {
    if (!it.split) continue;
    const BlockSpec& b = *it.spec;
    const double fp64 = ctx.gemm_mode() == GemmModeAuto
                            ? fp64_estimate_seconds(ctx, cplx, opA, SPLA_OP_NONE, b.m, b.n, b.k,
                                                    /*wholeChip=*/false)
                            : 0.0;
    it.split = it.p->plan(fp64);
    if (it.split) laneBytes[it.lane] = std::max(laneBytes[it.lane], it.p->workspace_bytes());
  }
  std::vector<Block> ws(nLanes);
  for (int i = 0; i < nLanes; ++i)
    if (laneBytes[i] > 0) ws[i] = Block(ctx.gpu_memory(), laneBytes[i]);
  std::vector<Block> reuse(nLanes);
  for (Item& it : items) {
    const BlockSpec& b = *it.spec;
    if (it.split) {
      it.p->enqueue_rest(lanes[it.lane], ws[it.lane].get(), keep);
    } else {
      FillSpec fill;
      fill.mode = b.fillMode;
      fill.rowOff = b.rowOffset;
      fill.colOff = b.colOffset;
      enqueue_on_device<T>(opA, SPLA_OP_NONE, b.m, b.n, b.k, alpha, static_cast<const T*>(b.A),
                           b.lda, static_cast<const T*>(b.B), b.ldb, beta, static_cast<T*>(b.C),
                           b.ldc, ctx, fill, lanes[it.lane], keep, &reuse[it.lane], false);
    }
  }
  if (!items.empty()) ctx.last_split() = items.back().p->report();
  // workspaces stay alive until the caller has drained the lanes
  for (Block& w : ws) keep.push_back(std::move(w));
  for (Block& r : reuse) keep.push_back(std::move(r));
}

}  // namespace

BatchPlan::BatchPlan(std::shared_ptr<spla::ContextInternal> ctx, bool complex, SplaOperation opA,
                     std::vector<BlockSpec> blocks, const double* alpha, const double* beta,
                     bool measure)
    : ctx_(std::move(ctx)), complex_(complex), opA_(opA), specs_(std::move(blocks)) {
  if (opA != SPLA_OP_TRANSPOSE && opA != SPLA_OP_CONJ_TRANSPOSE)
    throw spla::InvalidParameterError();
  alpha_[0] = alpha[0];
  alpha_[1] = complex ? alpha[1] : 0.0;
  beta_[0] = beta[0];
  beta_[1] = complex ? beta[1] : 0.0;
  // parameter and pointer checks precede any device work
  for (const BlockSpec& b : specs_) {
    if (b.m < 0 || b.n < 0 || b.k < 0 || b.rowOffset < 0 || b.colOffset < 0)
      throw spla::InvalidParameterError();
    if (b.fillMode != SPLA_FILL_MODE_FULL && b.fillMode != SPLA_FILL_MODE_UPPER &&
        b.fillMode != SPLA_FILL_MODE_LOWER)
      throw spla::InvalidParameterError();
    if (b.m == 0 || b.n == 0) continue;
    if (b.lda < std::max(1, b.k) || b.ldb < std::max(1, b.k) || b.ldc < b.m)
      throw spla::InvalidParameterError();
    if ((b.k > 0 && (!b.A || !b.B)) || !b.C) throw spla::InvalidPointerError();
  }
  DeviceGuard guard(ctx_->device());
  for (const BlockSpec& b : specs_) {
    if (b.m == 0 || b.n == 0) continue;
    if ((b.k > 0 && (!on_device(b.A) || !on_device(b.B))) || !on_device(b.C))
      throw spla::GPUInvalidDevicePointerError();
  }
  BatchCandidate best = build(TileL, 1.0);
  BatchCandidate other = build(TileM, 1.0);
  if (other.modelSeconds < best.modelSeconds) std::swap(best, other);
  if (!kPersistentLaunch) {
    // switched off: the blocks run one by one on the single-product engine
    cand_ = std::move(best);
    return;
  }
  if (measure && !best.items.empty()) {
    std::vector<BatchCandidate> cands;
    for (int cfg : {TileL, TileM})
      for (double sc : {0.5, 1.0, 2.0}) cands.push_back(build(cfg, sc));
    double bestMs = 1e300;
    for (BatchCandidate& c : cands) {
      const double ms = time_candidate(c);
      if (ms < bestMs) {
        bestMs = ms;
        best = c;
      }
    }
    measuredMs_ = bestMs;
  }
  cand_ = std::move(best);
  upload(cand_);
  capture();
}

BatchPlan::~BatchPlan() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
}

BatchCandidate BatchPlan::build(int cfg, double splitScale) const {
  BatchCandidate c;
  c.cfg = cfg;
  c.splitScale = splitScale;
  const TileShape ts = tile_shape(cfg);
  const int comps = complex_ ? 2 : 1;
  const int ar = ts.bm / comps;  // element rows per tile
  const int slots = ctx_->num_cus() * blocks_per_cu(cfg);
  const double slotRate = kPeakPerCU * efficiency(cfg) / blocks_per_cu(cfg);

  struct Tile {
    int block, tm, tn;
    double cost;
  };
  std::vector<Tile> tiles;
  std::vector<int> tilesM(specs_.size(), 0), tilesN(specs_.size(), 0);
  double total = 0.0;
  for (std::size_t b = 0; b < specs_.size(); ++b) {
    const BlockSpec& s = specs_[b];
    if (s.m == 0 || s.n == 0) continue;
    const int64_t Kp = static_cast<int64_t>(s.k) * comps;
    tilesM[b] = (s.m + ar - 1) / ar;
    tilesN[b] = (s.n + ts.bn - 1) / ts.bn;
    const double cost = 2.0 * ts.bm * ts.bn * static_cast<double>(std::max<int64_t>(Kp, ts.bk)) /
                        slotRate;
    for (int tm = 0; tm < tilesM[b]; ++tm)
      for (int tn = 0; tn < tilesN[b]; ++tn) {
        const int64_t r = static_cast<int64_t>(tm) * ar, col = static_cast<int64_t>(tn) * ts.bn;
        if (!tile_active(s.fillMode, s.rowOffset, s.colOffset, r, col,
                         std::min<int64_t>(ar, s.m - r), std::min<int64_t>(ts.bn, s.n - col)))
          continue;
        tiles.push_back({static_cast<int>(b), tm, tn, cost});
        total += cost;
      }
  }
  // target cost of one item: at least three items per slot, at least ~20 us
  const double target = std::max(total / (3.0 * std::max(1, slots)), 20e-6);

  c.blocks.resize(specs_.size());
  std::vector<int> splits(specs_.size(), 1);
  for (std::size_t b = 0; b < specs_.size(); ++b) {
    const BlockSpec& s = specs_[b];
    BatchBlock& bb = c.blocks[b];
    std::memset(&bb, 0, sizeof(bb));
    if (s.m == 0 || s.n == 0) continue;
    const int64_t Kp = static_cast<int64_t>(s.k) * comps;
    const int64_t kBlocks = (Kp + ts.bk - 1) / ts.bk;
    const double tileCost = 2.0 * ts.bm * ts.bn * static_cast<double>(std::max<int64_t>(Kp, 1)) /
                            slotRate;
    int64_t sp = static_cast<int64_t>(std::ceil(tileCost * splitScale / target));
    sp = std::max<int64_t>(1, std::min<int64_t>(sp, std::max<int64_t>(1, kBlocks / 16)));
    const int64_t kbPer = kBlocks > 0 ? (kBlocks + sp - 1) / sp : 1;
    sp = kBlocks > 0 ? (kBlocks + kbPer - 1) / kbPer : 1;
    splits[b] = static_cast<int>(sp);
    bb.A = static_cast<const double*>(s.A);
    bb.B = static_cast<const double*>(s.B);
    bb.C = static_cast<double*>(s.C);
    bb.lda = s.lda;
    bb.ldb = s.ldb;
    bb.ldc = s.ldc;
    bb.m = s.m;
    bb.n = s.n;
    bb.Kp = Kp;
    bb.Mp = s.m * comps;
    bb.tilesN = tilesN[b];
    bb.splits = static_cast<int>(sp);
    bb.kSplit = kbPer * ts.bk;
    if (sp > 1) {
      bb.workOffset = c.workDoubles;
      c.workDoubles += sp * static_cast<int64_t>(s.n) * bb.Mp;
      bb.counterBase = c.counters;
      c.counters += tilesM[b] * tilesN[b];
    }
  }
  // items in decreasing cost order; LPT simulation of the persistent schedule
  struct Costed {
    BatchItem item;
    double cost;
  };
  std::vector<Costed> items;
  for (const Tile& t : tiles) {
    const int sp = splits[t.block];
    // the last split of a tile also adds up the partial sums of all splits
    const double reduce =
        sp > 1 ? (sp + 2.0) * ar * ts.bn * 8.0 * comps / kWorkgroupBytesPerSecond : 0.0;
    for (int q = 0; q < sp; ++q)
      items.push_back({{t.block, t.tm, t.tn, q}, t.cost / sp + kItemOverhead + reduce / sp});
  }
  std::stable_sort(items.begin(), items.end(),
                   [](const Costed& a, const Costed& b) { return a.cost > b.cost; });
  c.items.reserve(items.size());
  for (const Costed& it : items) c.items.push_back(it.item);
  c.workgroups = static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(slots, items.size())));
  std::priority_queue<double, std::vector<double>, std::greater<double>> load;
  for (int i = 0; i < c.workgroups; ++i) load.push(0.0);
  double makespan = 0.0;
  for (const Costed& it : items) {
    double l = load.top();
    load.pop();
    l += it.cost;
    makespan = std::max(makespan, l);
    load.push(l);
  }
  c.modelSeconds = makespan + kLaunchOverhead;
  return c;
}

void BatchPlan::upload(const BatchCandidate& c, std::vector<BatchBlock>* override) {
  const hipStream_t s = ctx_->compute_stream();
  const std::vector<BatchBlock>& blocks = override ? *override : c.blocks;
  dBlocks_ = Block(ctx_->gpu_memory(), std::max<std::size_t>(1, blocks.size()) * sizeof(BatchBlock));
  dItems_ = Block(ctx_->gpu_memory(), std::max<std::size_t>(1, c.items.size()) * sizeof(BatchItem));
  dWork_ = Block(ctx_->gpu_memory(), std::max<int64_t>(1, c.workDoubles) * sizeof(double));
  dCounters_ = Block(ctx_->gpu_memory(), (c.counters + 2) * sizeof(unsigned));
  if (!blocks.empty())
    hip_check(hipMemcpyAsync(dBlocks_.get(), blocks.data(), blocks.size() * sizeof(BatchBlock),
                             hipMemcpyHostToDevice, s));
  if (!c.items.empty())
    hip_check(hipMemcpyAsync(dItems_.get(), c.items.data(), c.items.size() * sizeof(BatchItem),
                             hipMemcpyHostToDevice, s));
  // counters and queue start at zero; every launch leaves them at zero
  hip_check(hipMemsetAsync(dCounters_.get(), 0, dCounters_.bytes(), s));
  hip_check(hipStreamSynchronize(s));
  std::memset(&desc_, 0, sizeof(desc_));
  desc_.blocks = dBlocks_.get<BatchBlock>();
  desc_.items = dItems_.get<BatchItem>();
  desc_.numItems = static_cast<int>(c.items.size());
  desc_.work = dWork_.get<double>();
  desc_.counters = dCounters_.get<unsigned>() + 2;
  desc_.queue = dCounters_.get<unsigned>();
  desc_.alphaRe = alpha_[0];
  desc_.alphaIm = alpha_[1];
  desc_.betaRe = beta_[0];
  desc_.betaIm = beta_[1];
  desc_.betaZero = (beta_[0] == 0.0 && beta_[1] == 0.0) ? 1 : 0;
  desc_.conjA = (complex_ && opA_ == SPLA_OP_CONJ_TRANSPOSE) ? 1 : 0;
  desc_.conjB = 0;
}

void BatchPlan::launch(hipStream_t s) {
  if (desc_.numItems == 0) return;
  hip_check(launch_batched_ssb(desc_, complex_, cand_.cfg, cand_.workgroups, s));
}

double BatchPlan::time_candidate(const BatchCandidate& c) {
  // Run the candidate on scratch copies of C (C itself is left untouched).
  const int comps = complex_ ? 2 : 1;
  int64_t total = 0;
  for (const BlockSpec& s : specs_) total += static_cast<int64_t>(s.m) * s.n * comps;
  Block scratch(ctx_->gpu_memory(), std::max<int64_t>(1, total) * sizeof(double));
  const hipStream_t st = ctx_->compute_stream();
  hip_check(hipMemsetAsync(scratch.get(), 0, scratch.bytes(), st));
  std::vector<BatchBlock> shadow = c.blocks;
  int64_t off = 0;
  for (std::size_t b = 0; b < specs_.size(); ++b) {
    shadow[b].C = scratch.get<double>() + off;
    shadow[b].ldc = specs_[b].m;
    off += static_cast<int64_t>(specs_[b].m) * specs_[b].n * comps;
  }
  const BatchCandidate saved = cand_;
  cand_ = c;
  upload(c, &shadow);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0));
  hip_check(hipEventCreate(&e1));
  launch(st);  // warm-up
  hip_check(hipEventRecord(e0, st));
  constexpr int kReps = 3;
  for (int r = 0; r < kReps; ++r) launch(st);
  hip_check(hipEventRecord(e1, st));
  hip_check(hipEventSynchronize(e1));
  float ms = 0.0f;
  hip_check(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  cand_ = saved;
  return ms / kReps;
}

void BatchPlan::capture() {
  if (desc_.numItems == 0) return;
  const hipStream_t s = ctx_->compute_stream();
  hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const hipError_t launched = launch_batched_ssb(desc_, complex_, cand_.cfg, cand_.workgroups, s);
  hipGraph_t g = nullptr;
  const hipError_t ended = hipStreamEndCapture(s, &g);
  hip_check(launched);
  hip_check(ended);
  graph_ = g;
  hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

void BatchPlan::execute() {
  DeviceGuard guard(ctx_->device());
  hip_check(hipStreamSynchronize(nullptr));
  const hipStream_t s = ctx_->compute_stream();
  ctx_->timer().reset();
  if (!kPersistentLaunch || ctx_->gemm_mode() != GemmModeFP64) {
    // Block by block on the single-product engine. FP64: independent blocks go
    // round robin to the compute stream and the context's batch streams, so
    // that small products run side by side, with one synchronization. Split
    // precision: the same lanes, staged so that all blocks share one host
    // synchronization for their plans (execute_split_batch).
    const bool split = ctx_->gemm_mode() != GemmModeFP64;
    constexpr int kLanes = 1 + spla::ContextInternal::kBatchStreams;
    hipStream_t lanes[kLanes] = {s};
    int nLanes = 1;
    hipEvent_t start;
    hip_check(hipEventCreateWithFlags(&start, hipEventDisableTiming));
    hip_check(hipEventRecord(start, s));
    for (int i = 0; i < spla::ContextInternal::kBatchStreams; ++i) {
      lanes[nLanes] = ctx_->batch_stream(i);
      hip_check(hipStreamWaitEvent(lanes[nLanes++], start, 0));
    }
    (void)hipEventDestroy(start);
    std::vector<Block> keep;
    if (split) {
      if (complex_) {
        using Z = std::complex<double>;
        execute_split_batch<Z>(*ctx_, specs_, opA_, Z(alpha_[0], alpha_[1]),
                               Z(beta_[0], beta_[1]), lanes, nLanes, keep);
      } else {
        execute_split_batch<double>(*ctx_, specs_, opA_, alpha_[0], beta_[0], lanes, nLanes, keep);
      }
    } else {
      Block reuse[kLanes];
      int next = 0;
      for (const BlockSpec& b : specs_) {
        if (b.m == 0 || b.n == 0) continue;
        const int lane = next;
        next = (next + 1) % nLanes;
        const hipStream_t ls = lanes[lane];
        FillSpec fill;
        fill.mode = b.fillMode;
        fill.rowOff = b.rowOffset;
        fill.colOff = b.colOffset;
        if (complex_) {
          using Z = std::complex<double>;
          enqueue_on_device<Z>(opA_, SPLA_OP_NONE, b.m, b.n, b.k, Z(alpha_[0], alpha_[1]),
                               static_cast<const Z*>(b.A), b.lda, static_cast<const Z*>(b.B),
                               b.ldb, Z(beta_[0], beta_[1]), static_cast<Z*>(b.C), b.ldc, *ctx_,
                               fill, ls, keep, &reuse[lane]);
        } else {
          enqueue_on_device<double>(opA_, SPLA_OP_NONE, b.m, b.n, b.k, alpha_[0],
                                    static_cast<const double*>(b.A), b.lda,
                                    static_cast<const double*>(b.B), b.ldb, beta_[0],
                                    static_cast<double*>(b.C), b.ldc, *ctx_, fill, ls, keep,
                                    &reuse[lane]);
        }
      }
      for (Block& r : reuse) keep.push_back(std::move(r));
    }
    for (int i = 1; i < nLanes; ++i) {  // the compute stream joins the batch streams
      hipEvent_t done;
      hip_check(hipEventCreateWithFlags(&done, hipEventDisableTiming));
      hip_check(hipEventRecord(done, lanes[i]));
      hip_check(hipStreamWaitEvent(s, done, 0));
      (void)hipEventDestroy(done);
    }
    hip_check(hipStreamSynchronize(s));
    if (ctx_->timer().enabled()) ctx_->timer().resolve();
    return;
  }
  if (exec_) {
    PhaseScope ps(ctx_->timer(), PhaseGemmF64, s);
    hip_check(hipGraphLaunch(exec_, s));
  }
  hip_check(hipStreamSynchronize(s));
  if (ctx_->timer().enabled()) ctx_->timer().resolve();
}

}  // namespace spla_amd
