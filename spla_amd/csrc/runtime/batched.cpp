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
#include "runtime/batch_schedule.hpp"
#include "runtime/engine.hpp"
#include "runtime/hip_check.hpp"
#include "runtime/split_gemm.hpp"
#include "spla/exceptions.hpp"

namespace spla_amd {
namespace {

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

// ---------------------------------------------------------------------------
// GroupedLaunch

GroupedLaunch::~GroupedLaunch() { release_graph(); }

void GroupedLaunch::release_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

void GroupedLaunch::build(spla::ContextInternal& ctx, bool complex, SplaOperation opA,
                          const std::vector<BlockSpec>& specs, const double* alpha,
                          const double* beta, const BatchSchedule& sched, bool persistent,
                          bool withGraph) {
  release_graph();
  const hipStream_t s = ctx.compute_stream();
  std::vector<BatchBlock> blocks = sched.blocks;
  for (std::size_t b = 0; b < specs.size(); ++b) {
    const BlockSpec& sp = specs[b];
    BatchBlock& bb = blocks[b];
    bb.A = static_cast<const double*>(sp.A);
    bb.B = static_cast<const double*>(sp.B);
    bb.C = static_cast<double*>(sp.C);
    bb.lda = sp.lda;
    bb.ldb = sp.ldb;
    bb.ldc = sp.ldc;
  }
  dBlocks_ = Block(ctx.gpu_memory(), std::max<std::size_t>(1, blocks.size()) * sizeof(BatchBlock));
  dItems_ = Block(ctx.gpu_memory(), std::max<std::size_t>(1, sched.items.size()) * sizeof(BatchItem));
  dReduce_ = Block(ctx.gpu_memory(),
                   std::max<std::size_t>(1, sched.reduce.size()) * sizeof(BatchReduceTile));
  dWork_ = Block(ctx.gpu_memory(), std::max<int64_t>(1, sched.workDoubles) * sizeof(double));
  dSlots_ = Block(ctx.gpu_memory(),
                  (sched.slotStart.size() + std::max<std::size_t>(1, sched.slotItems.size())) *
                      sizeof(int));
  if (!blocks.empty())
    hip_check(hipMemcpyAsync(dBlocks_.get(), blocks.data(), blocks.size() * sizeof(BatchBlock),
                             hipMemcpyHostToDevice, s));
  if (!sched.items.empty())
    hip_check(hipMemcpyAsync(dItems_.get(), sched.items.data(),
                             sched.items.size() * sizeof(BatchItem), hipMemcpyHostToDevice, s));
  if (!sched.reduce.empty())
    hip_check(hipMemcpyAsync(dReduce_.get(), sched.reduce.data(),
                             sched.reduce.size() * sizeof(BatchReduceTile), hipMemcpyHostToDevice,
                             s));
  if (!sched.slotStart.empty())
    hip_check(hipMemcpyAsync(dSlots_.get(), sched.slotStart.data(),
                             sched.slotStart.size() * sizeof(int), hipMemcpyHostToDevice, s));
  if (!sched.slotItems.empty())
    hip_check(hipMemcpyAsync(dSlots_.get<int>() + sched.slotStart.size(), sched.slotItems.data(),
                             sched.slotItems.size() * sizeof(int), hipMemcpyHostToDevice, s));
  hip_check(hipStreamSynchronize(s));
  std::memset(&desc_, 0, sizeof(desc_));
  desc_.blocks = dBlocks_.get<BatchBlock>();
  desc_.items = dItems_.get<BatchItem>();
  desc_.slotStart = dSlots_.get<int>();
  desc_.slotItems = dSlots_.get<int>() + sched.slotStart.size();
  desc_.reduce = dReduce_.get<BatchReduceTile>();
  desc_.work = dWork_.get<double>();
  desc_.alphaRe = alpha[0];
  desc_.alphaIm = complex ? alpha[1] : 0.0;
  desc_.betaRe = beta[0];
  desc_.betaIm = complex ? beta[1] : 0.0;
  desc_.betaZero = (desc_.betaRe == 0.0 && desc_.betaIm == 0.0) ? 1 : 0;
  desc_.conjA = (complex && opA == SPLA_OP_CONJ_TRANSPOSE) ? 1 : 0;
  desc_.conjB = 0;
  kernels_ = batched_kernels(complex, opA == SPLA_OP_NONE ? MNMajor : KMajor, sched.cfg,
                             persistent);
  // GEMM workgroups: one per item (grouped) or one per slot (persistent)
  numItems_ = sched.items.empty() ? 0u
                                  : static_cast<unsigned>(persistent ? sched.slotStart.size() - 1
                                                                     : sched.items.size());
  numReduce_ = static_cast<unsigned>(sched.reduce.size());
  persistent_ = persistent;
  if (withGraph && numItems_ > 0) {
    // the two launches as the nodes of one graph, replayed by run()
    hip_check(hipGraphCreate(&graph_, 0));
    void* args[] = {&desc_};
    hipKernelNodeParams p;
    std::memset(&p, 0, sizeof(p));
    p.func = const_cast<void*>(kernels_.gemm);
    p.gridDim = dim3(numItems_);
    p.blockDim = dim3(kernels_.threads);
    p.sharedMemBytes = 0;
    p.kernelParams = args;
    p.extra = nullptr;
    hipGraphNode_t gemmNode = nullptr;
    hip_check(hipGraphAddKernelNode(&gemmNode, graph_, nullptr, 0, &p));
    if (numReduce_ > 0) {
      p.func = const_cast<void*>(kernels_.reduce);
      p.gridDim = dim3(numReduce_);
      hipGraphNode_t reduceNode = nullptr;
      hip_check(hipGraphAddKernelNode(&reduceNode, graph_, &gemmNode, 1, &p));
    }
    hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
}

void GroupedLaunch::launch(hipStream_t s) {
  void* args[] = {&desc_};
  if (numItems_ > 0)
    hip_check(hipLaunchKernel(kernels_.gemm, dim3(numItems_), dim3(kernels_.threads), args, 0, s));
  if (numReduce_ > 0)
    hip_check(
        hipLaunchKernel(kernels_.reduce, dim3(numReduce_), dim3(kernels_.threads), args, 0, s));
}

void GroupedLaunch::run(hipStream_t s) {
  if (exec_)
    hip_check(hipGraphLaunch(exec_, s));
  else
    launch(s);
}

// ---------------------------------------------------------------------------
// BatchPlan

BatchPlan::BatchPlan(std::shared_ptr<spla::ContextInternal> ctx, bool complex, SplaOperation opA,
                     std::vector<BlockSpec> blocks, const double* alpha, const double* beta,
                     int tuning)
    : ctx_(std::move(ctx)), complex_(complex), opA_(opA), specs_(std::move(blocks)) {
  const bool measure = tuning == 1;
  // ssb blocks: op(A) = A^T or A^H (A k x m); sbs blocks: op(A) = A (A m x k, full fill)
  if (opA != SPLA_OP_TRANSPOSE && opA != SPLA_OP_CONJ_TRANSPOSE && opA != SPLA_OP_NONE)
    throw spla::InvalidParameterError();
  const bool sbs = opA == SPLA_OP_NONE;
  alpha_[0] = alpha[0];
  alpha_[1] = complex ? alpha[1] : 0.0;
  beta_[0] = beta[0];
  beta_[1] = complex ? beta[1] : 0.0;
  // parameter and pointer checks precede any device work
  for (const BlockSpec& b : specs_) {
    if (b.m < 0 || b.n < 0 || b.k < 0 || b.rowOffset < 0 || b.colOffset < 0)
      throw spla::InvalidParameterError();
    if (b.fillMode != SPLA_FILL_MODE_FULL &&
        (sbs || (b.fillMode != SPLA_FILL_MODE_UPPER && b.fillMode != SPLA_FILL_MODE_LOWER)))
      throw spla::InvalidParameterError();
    if (b.m == 0 || b.n == 0) continue;
    if (b.lda < std::max(1, sbs ? b.m : b.k) || b.ldb < std::max(1, b.k) || b.ldc < b.m)
      throw spla::InvalidParameterError();
    if ((b.k > 0 && (!b.A || !b.B)) || !b.C) throw spla::InvalidPointerError();
  }
  DeviceGuard guard(ctx_->device());
  for (const BlockSpec& b : specs_) {
    if (b.m == 0 || b.n == 0) continue;
    if ((b.k > 0 && (!on_device(b.A) || !on_device(b.B))) || !on_device(b.C))
      throw spla::GPUInvalidDevicePointerError();
  }
  // Autotuning: the model picks the tile shape (and scales the split counts);
  // in measured mode every candidate also runs on scratch copies of C, the
  // fastest is kept and its measured rate corrects the model's efficiency for
  // the batches the context plans afterwards.
  BatchSchedule best = best_schedule(specs_);
  bool persistent = tuning == 2 || ctx_->batch_persistent();
  if (measure && !best.items.empty()) {
    double bestMs = 1e300;
    for (int cfg : {TileL, TileM})
      for (double sc : {0.5, 1.0, 2.0}) {
        BatchSchedule c = schedule(specs_, cfg, sc);
        for (bool pers : {false, true}) {
          const double ms = time_candidate(c, pers);
          if (ms < bestMs) {
            bestMs = ms;
            best = c;
            persistent = pers;
          }
        }
      }
    measuredMs_ = bestMs;
    const double eff = ctx_->batch_efficiency(best.cfg) * best.modelSeconds / (bestMs * 1e-3);
    ctx_->set_batch_efficiency(best.cfg, std::min(0.95, std::max(0.05, eff)));
    ctx_->set_batch_persistent(persistent);
  }
  sched_ = std::move(best);
  full_.build(*ctx_, complex_, opA_, specs_, alpha_, beta_, sched_, persistent,
              /*withGraph=*/true);
}

BatchPlan::~BatchPlan() = default;

BatchSchedule BatchPlan::schedule(const std::vector<BlockSpec>& specs, int cfg,
                                  double splitScale) const {
  std::vector<BatchShape> shapes;
  shapes.reserve(specs.size());
  for (const BlockSpec& b : specs)
    shapes.push_back({b.m, b.n, b.k, b.fillMode, b.rowOffset, b.colOffset});
  const bool alphaZero = alpha_[0] == 0.0 && alpha_[1] == 0.0;
  return plan_batch_schedule(shapes, complex_, alphaZero, cfg, splitScale, ctx_->num_cus(),
                             ctx_->batch_efficiency(cfg));
}

BatchSchedule BatchPlan::best_schedule(const std::vector<BlockSpec>& specs) const {
  BatchSchedule best = schedule(specs, TileL, 1.0);
  BatchSchedule other = schedule(specs, TileM, 1.0);
  if (other.modelSeconds < best.modelSeconds) std::swap(best, other);
  return best;
}

double BatchPlan::time_candidate(const BatchSchedule& c, bool persistent) {
  // Run the candidate on scratch copies of C (C itself is left untouched).
  const int comps = complex_ ? 2 : 1;
  int64_t total = 0;
  for (const BlockSpec& s : specs_) total += static_cast<int64_t>(s.m) * s.n * comps;
  Block scratch(ctx_->gpu_memory(), std::max<int64_t>(1, total) * sizeof(double));
  const hipStream_t st = ctx_->compute_stream();
  hip_check(hipMemsetAsync(scratch.get(), 0, scratch.bytes(), st));
  std::vector<BlockSpec> shadow = specs_;
  int64_t off = 0;
  for (BlockSpec& b : shadow) {
    b.C = scratch.get<double>() + off;
    b.ldc = std::max(1, b.m);
    off += static_cast<int64_t>(b.m) * b.n * comps;
  }
  GroupedLaunch trial;
  trial.build(*ctx_, complex_, opA_, shadow, alpha_, beta_, c, persistent, /*withGraph=*/false);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0));
  hip_check(hipEventCreate(&e1));
  trial.launch(st);  // warm-up
  hip_check(hipEventRecord(e0, st));
  constexpr int kReps = 3;
  for (int r = 0; r < kReps; ++r) trial.launch(st);
  hip_check(hipEventRecord(e1, st));
  hip_check(hipEventSynchronize(e1));
  float ms = 0.0f;
  hip_check(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / kReps;
}

void BatchPlan::execute() {
  DeviceGuard guard(ctx_->device());
  hip_check(hipStreamSynchronize(nullptr));
  const hipStream_t s = ctx_->compute_stream();
  ctx_->timer().reset();
  const bool alphaZero = alpha_[0] == 0.0 && alpha_[1] == 0.0;
  const int mode = ctx_->gemm_mode();
  // Automatic mode: the blocks whose shape rules the split form out (a lower
  // bound of its cost with the fewest moduli any data could need, against the
  // block's own FP64 estimate) stay on FP64; the others are split candidates.
  // The candidates would otherwise join the grouped launch, so they run split
  // only when their lower bounds, one after another on the whole chip, beat
  // the grouped launch's model time for them.
  std::vector<BlockSpec> fp64, split;
  std::vector<char> mask;
  bool allFp64 = mode == GemmModeFP64 || alphaZero;
  if (!allFp64 && mode == GemmModeAuto) {
    mask.assign(specs_.size(), 0);
    double splitSeconds = 0.0;
    for (std::size_t i = 0; i < specs_.size(); ++i) {
      const BlockSpec& b = specs_[i];
      if (b.m == 0 || b.n == 0) continue;
      double lower = 0.0;
      mask[i] = may_split(b, &lower) ? 1 : 0;
      if (mask[i]) splitSeconds += lower;
      (mask[i] ? split : fp64).push_back(b);
    }
    allFp64 = split.empty() || splitSeconds >= 0.8 * best_schedule(split).modelSeconds;
  }
  if (allFp64) {
    // the grouped GEMM and the split reduction, replayed from the graph
    lastSplitBlocks_ = 0;
    PhaseScope ps(ctx_->timer(), PhaseGemmF64, s);
    full_.run(s);
  } else {
    // Split precision, or the automatic choice: the FP64 blocks run as one
    // grouped launch on the compute stream; the split blocks, and every block in
    // split mode, go block by block over the context's batch streams, staged so
    // that all of them share one host synchronization for their plans
    // (execute_split_batch; blocks whose plan falls back run the FP64 kernels).
    if (mode == GemmModeAuto) {
      if (!fp64.empty() && (!autoLaunch_ || mask != autoMask_)) {
        autoLaunch_ = std::make_unique<GroupedLaunch>();
        autoLaunch_->build(*ctx_, complex_, opA_, fp64, alpha_, beta_, best_schedule(fp64),
                           ctx_->batch_persistent(), /*withGraph=*/true);
        autoMask_ = mask;
      }
    } else {
      split = specs_;
    }
    lastSplitBlocks_ = static_cast<int>(std::count_if(
        split.begin(), split.end(), [](const BlockSpec& b) { return b.m > 0 && b.n > 0; }));
    // lanes of the block-by-block products: the batch streams, and the compute
    // stream too when no grouped launch occupies it
    constexpr int kMaxLanes = 1 + spla::ContextInternal::kBatchStreams;
    hipStream_t lanes[kMaxLanes];
    int nLanes = 0;
    if (fp64.empty()) lanes[nLanes++] = s;
    hipEvent_t start;
    hip_check(hipEventCreateWithFlags(&start, hipEventDisableTiming));
    hip_check(hipEventRecord(start, s));
    for (int i = 0; i < spla::ContextInternal::kBatchStreams; ++i) {
      lanes[nLanes] = ctx_->batch_stream(i);
      hip_check(hipStreamWaitEvent(lanes[nLanes++], start, 0));
    }
    (void)hipEventDestroy(start);
    if (!fp64.empty()) {
      PhaseScope ps(ctx_->timer(), PhaseGemmF64, s);
      autoLaunch_->run(s);
    }
    std::vector<Block> keep;
    if (!split.empty()) {
      if (complex_) {
        using Z = std::complex<double>;
        execute_split_batch<Z>(*ctx_, split, opA_, Z(alpha_[0], alpha_[1]),
                               Z(beta_[0], beta_[1]), lanes, nLanes, keep);
      } else {
        execute_split_batch<double>(*ctx_, split, opA_, alpha_[0], beta_[0], lanes, nLanes,
                                    keep);
      }
    }
    for (int i = 0; i < nLanes; ++i) {  // the compute stream joins the batch streams
      if (lanes[i] == s) continue;
      hipEvent_t done;
      hip_check(hipEventCreateWithFlags(&done, hipEventDisableTiming));
      hip_check(hipEventRecord(done, lanes[i]));
      hip_check(hipStreamWaitEvent(s, done, 0));
      (void)hipEventDestroy(done);
    }
  }
  hip_check(hipStreamSynchronize(s));
  if (ctx_->timer().enabled()) ctx_->timer().resolve();
}

bool BatchPlan::may_split(const BlockSpec& b, double* lowerBound) const {
  const double fp64 = fp64_estimate_seconds(*ctx_, complex_, opA_, SPLA_OP_NONE, b.m, b.n, b.k,
                                            /*wholeChip=*/false);
  if (complex_) {
    using Z = std::complex<double>;
    SplitProduct<Z> p(opA_, SPLA_OP_NONE, b.m, b.n, b.k, Z(alpha_[0], alpha_[1]),
                      static_cast<const Z*>(b.A), b.lda, static_cast<const Z*>(b.B), b.ldb,
                      Z(beta_[0], beta_[1]), static_cast<Z*>(b.C), b.ldc, *ctx_, FillSpec());
    *lowerBound = p.lower_bound_seconds();
    return p.may_beat_fp64(fp64);
  }
  SplitProduct<double> p(opA_, SPLA_OP_NONE, b.m, b.n, b.k, alpha_[0],
                         static_cast<const double*>(b.A), b.lda, static_cast<const double*>(b.B),
                         b.ldb, beta_[0], static_cast<double*>(b.C), b.ldc, *ctx_, FillSpec());
  *lowerBound = p.lower_bound_seconds();
  return p.may_beat_fp64(fp64);
}

}  // namespace spla_amd
