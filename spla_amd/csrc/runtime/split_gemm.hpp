// Split-precision (Ozaki scheme II) products on the INT8 matrix cores.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <vector>

#include "kernels/ozaki_desc.hpp"
#include "runtime/allocators.hpp"
#include "runtime/context_internal.hpp"
#include "runtime/engine.hpp"
#include "runtime/split_plan.hpp"
#include "runtime/split_report.hpp"
#include "spla/types.h"

namespace spla_amd {

// One product C <- alpha op(A) op(B) + beta C on device-resident operands in
// three stages, so that many products can share one host synchronization:
//   1. enqueue_stats: per-row statistics of both operands on the device,
//      copied to pinned host memory;
//   2. plan: after the stream has drained, the host plan (widths, moduli,
//      exponents) or the reason to fall back to the FP64 kernels;
//   3. enqueue_rest: residue planes, the batched INT8 GEMM and the
//      reconstruction with the alpha / beta update.
template <typename T>
class SplitProduct {
public:
  SplitProduct(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k, T alpha,
               const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc,
               spla::ContextInternal& ctx, const FillSpec& fill);

  // False (nothing enqueued) when the split form does not apply: FP64 mode,
  // an empty product or a reduction longer than the residue rows allow.
  bool enqueue_stats(hipStream_t stream);
  // Automatic mode, before any statistics: false (report().reason =
  // SplitNotFaster) when even the fewest moduli the tolerance allows, at the
  // full-chip INT8 rate, cannot make the split form faster than fp64Seconds.
  // plan() would reject every such product, so its statistics pass is skipped.
  bool may_beat_fp64(double fp64Seconds);
  // The cost may_beat_fp64 compares: a lower bound of the split form's time
  // (fewest moduli, full-chip INT8 rate, no tile padding, one reduction split).
  double lower_bound_seconds() const;
  // False when the product falls back to the FP64 kernels (report().reason).
  bool plan(double fp64Seconds);
  // Device bytes of the residue planes and INT8 GEMM output (after plan()).
  std::size_t workspace_bytes() const;
  // Longest reduction chunk (a multiple of 256) whose workspace, at this plan's
  // moduli, fits `limit` bytes; 0 when none does (after plan()).
  int64_t max_chunk_k(std::size_t limit) const;
  // `workspace`: at least workspace_bytes(), 256-byte aligned, or nullptr to
  // allocate it into `keep`; the other buffers are appended to `keep`, which
  // must live until the stream has drained.
  void enqueue_rest(hipStream_t stream, void* workspace, std::vector<Block>& keep);

  const SplitReport& report() const { return rep_; }

private:
  SplaOperation opA_, opB_;
  int64_t m_, n_, k_;
  T alpha_, beta_;
  const T* A_;
  const T* B_;
  T* C_;
  int64_t lda_, ldb_, ldc_;
  spla::ContextInternal& ctx_;
  FillSpec fill_;
  int la_ = 0, lb_ = 0;
  Block dStat_, hStat_;
  SplitPlan plan_;
  SplitReport rep_;
  int G_ = 0;
  int64_t kpad_ = 0, mpad_ = 0, npad_ = 0;
  int cfg_ = 0, bm_ = 0, bn_ = 0;
  int64_t splits_ = 1, kSplit_ = 0;
};

// The single-product form: the three stages with one synchronization of
// `stream` in between. Returns true when the product was enqueued (in split
// precision, or in reduction chunks of which some may run on the FP64 kernels),
// false (nothing but the statistics enqueued) when the product falls back to the
// FP64 kernels. The report is stored in the context.
//
// The workspace (residue planes and GEMM output) is kept within the context's
// split workspace limit: a product whose workspace exceeds it runs in reduction
// chunks of at most the limit, one after the other in one workspace; a product
// that cannot be chunked (the limit holds less than 4096 reduction indices) or
// whose workspace cannot be allocated falls back (reason SplitNoWorkspace).
// `reuse` (optional): a workspace block shared by the products of one stream,
// grown when too small, so that the streamed k-chunks of host-resident operands
// hold one workspace between them.
template <typename T>
bool enqueue_split_product(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                           T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                           T* C, int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill,
                           double fp64Seconds, hipStream_t stream, std::vector<Block>& keep,
                           Block* reuse = nullptr);

}  // namespace spla_amd
