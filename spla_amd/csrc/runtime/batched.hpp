// Batched ring-step block products: many independent
// C_b <- alpha op(A_b) B_b + beta C_b, the block products of one pgemm_ssb /
// pgemm_ssbtr ring step (op(A) = A^T or A^H, src/pgemm_ssb/ring_ssb_gpu.cpp:193-197)
// or of one pgemm_sbs ring step (op(A) = A, src/pgemm_sbs/ring_sbs_gpu.cpp:57-88).
//
// FP64 (and alpha == 0): one GEMM launch over the tiles and reduction splits of
// every block (grouped: a workgroup per item; or persistent: a workgroup per
// resident slot with a static list of items) and one fixed-order reduction
// launch over the split tiles, both nodes of one hipGraph built at planning
// time and replayed by every execute(). Split precision runs block by block on
// the single-product engine over four streams; the automatic mode sends the
// blocks the split form cannot win to a grouped FP64 launch and the others
// block by block.
//
// Autotuning: for each tile shape the schedule (batch_schedule.hpp) picks the
// per-block split counts from a cost model of the launch; the cheaper shape
// wins. In measured mode twelve candidates (two tile shapes x three split
// scales x grouped / persistent dispatch) also run on scratch copies of C, the
// fastest is kept, and its measured time corrects the context's efficiency of
// that tile shape (and its dispatch becomes the context's choice) for the
// batches it plans later.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "kernels/batched_desc.hpp"
#include "kernels/launch.hpp"
#include "runtime/allocators.hpp"
#include "runtime/batch_schedule.hpp"
#include "runtime/context_internal.hpp"
#include "spla/types.h"

namespace spla_amd {

struct BlockSpec {
  int m, n, k;
  const void* A;
  int lda;
  const void* B;
  int ldb;
  void* C;
  int ldc;
  int fillMode;
  int rowOffset, colOffset;
};

// One grouped launch (GEMM over the items, then the split reduction) over a
// list of blocks: device copies of a schedule and, optionally, the hipGraph
// holding both launches.
class GroupedLaunch {
public:
  GroupedLaunch() = default;
  ~GroupedLaunch();
  GroupedLaunch(const GroupedLaunch&) = delete;
  GroupedLaunch& operator=(const GroupedLaunch&) = delete;

  // persistent: one workgroup per slot of the schedule's static lists, else one
  // per item
  void build(spla::ContextInternal& ctx, bool complex, SplaOperation opA,
             const std::vector<BlockSpec>& specs, const double* alpha, const double* beta,
             const BatchSchedule& sched, bool persistent, bool withGraph);
  void run(hipStream_t s);     // the graph (or the two launches)
  void launch(hipStream_t s);  // the two launches, without the graph
  unsigned workgroups() const { return numItems_ + numReduce_; }
  bool persistent() const { return persistent_; }

private:
  void release_graph();

  Block dBlocks_, dItems_, dReduce_, dWork_, dSlots_;
  BatchDesc desc_{};
  BatchKernels kernels_{};
  unsigned numItems_ = 0, numReduce_ = 0;
  bool persistent_ = false;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
};

class BatchPlan {
public:
  // tuning: 0 the model's schedule with the context's dispatch, 1 measured
  // candidates, 2 the model's schedule with persistent dispatch
  BatchPlan(std::shared_ptr<spla::ContextInternal> ctx, bool complex, SplaOperation opA,
            std::vector<BlockSpec> blocks, const double* alpha, const double* beta, int tuning);
  ~BatchPlan();
  BatchPlan(const BatchPlan&) = delete;
  BatchPlan& operator=(const BatchPlan&) = delete;

  void execute();

  int tile_config() const { return sched_.cfg; }
  int num_items() const { return static_cast<int>(sched_.items.size()); }
  // workgroups of the GEMM launch (one per item) plus those of the reduction
  int workgroups() const { return static_cast<int>(full_.workgroups()); }
  double model_ms() const { return sched_.modelSeconds * 1e3; }
  double measured_ms() const { return measuredMs_; }
  // Blocks the last execute() ran in the split form, block by block (0: every
  // block in the grouped FP64 launch).
  int last_split_blocks() const { return lastSplitBlocks_; }
  bool persistent() const { return full_.persistent(); }

private:
  BatchSchedule schedule(const std::vector<BlockSpec>& specs, int cfg, double splitScale) const;
  BatchSchedule best_schedule(const std::vector<BlockSpec>& specs) const;
  double time_candidate(const BatchSchedule& c, bool persistent);
  // automatic mode: whether the split form could beat FP64 on this block's shape
  // Whether block b may run faster split than on FP64 (its own estimates), and
  // the lower bound of its split-form time.
  bool may_split(const BlockSpec& b, double* lowerBound) const;

  std::shared_ptr<spla::ContextInternal> ctx_;
  bool complex_;
  SplaOperation opA_;
  std::vector<BlockSpec> specs_;
  double alpha_[2], beta_[2];
  BatchSchedule sched_;
  GroupedLaunch full_;
  // automatic mode: the grouped launch of the blocks that stay on FP64
  int lastSplitBlocks_ = 0;
  std::unique_ptr<GroupedLaunch> autoLaunch_;
  std::vector<char> autoMask_;
  double measuredMs_ = 0.0;
};

}  // namespace spla_amd
