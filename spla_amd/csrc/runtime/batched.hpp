// Batched ring-step block products: many independent
// C_b <- alpha op(A_b) B_b + beta C_b (op = transpose or conjugate transpose,
// the block products of one pgemm_ssb / pgemm_ssbtr ring step,
// src/pgemm_ssb/ring_ssb_gpu.cpp:193-197) executed by one persistent launch
// that is captured once in a hipGraph and replayed.
//
// Autotuning: for each tile shape the plan picks per-block split factors from
// the target item cost and simulates the persistent schedule (items in
// decreasing cost order on the CU slots, LPT); the cheaper shape wins. In
// measured mode the candidate plans are also run on scratch copies of C and
// the fastest is kept.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "kernels/batched_desc.hpp"
#include "runtime/allocators.hpp"
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

struct BatchCandidate {
  int cfg = 0;
  double splitScale = 1.0;
  std::vector<BatchBlock> blocks;
  std::vector<BatchItem> items;
  int64_t workDoubles = 0;
  int counters = 0;
  int workgroups = 0;
  double modelSeconds = 0.0;
};

class BatchPlan {
public:
  BatchPlan(std::shared_ptr<spla::ContextInternal> ctx, bool complex, SplaOperation opA,
            std::vector<BlockSpec> blocks, const double* alpha, const double* beta, bool measure);
  ~BatchPlan();
  BatchPlan(const BatchPlan&) = delete;
  BatchPlan& operator=(const BatchPlan&) = delete;

  void execute();

  int tile_config() const { return cand_.cfg; }
  int num_items() const { return static_cast<int>(cand_.items.size()); }
  int workgroups() const { return cand_.workgroups; }
  double model_ms() const { return cand_.modelSeconds * 1e3; }
  double measured_ms() const { return measuredMs_; }

private:
  BatchCandidate build(int cfg, double splitScale) const;
  void upload(const BatchCandidate& c, std::vector<BatchBlock>* override = nullptr);
  void launch(hipStream_t s);
  double time_candidate(const BatchCandidate& c);
  void capture();

  std::shared_ptr<spla::ContextInternal> ctx_;
  bool complex_;
  SplaOperation opA_;
  std::vector<BlockSpec> specs_;
  double alpha_[2], beta_[2];
  BatchCandidate cand_;
  Block dBlocks_, dItems_, dWork_, dCounters_;
  BatchDesc desc_{};
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  double measuredMs_ = 0.0;
};

}  // namespace spla_amd
