// Entry points of include/spla_amd/spla_amd.h (no SPLA counterpart).
#include <algorithm>
#include <cstddef>

#include "api/internal_access.hpp"
#include "runtime/batched.hpp"
#include "runtime/ring_plan.hpp"
#include "spla_amd/spla_amd.h"

extern "C" {

SPLA_EXPORT SplaError spla_amd_ctx_set_stream(SplaContext ctx, void* hipStream) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx))
                   .set_user_stream(reinterpret_cast<hipStream_t>(hipStream)))
}

SPLA_EXPORT SplaError spla_amd_ctx_set_stream_chunk_bytes(SplaContext ctx, size_t bytes) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  if (bytes == 0) return SPLA_INVALID_PARAMETER_ERROR;
  SPLA_C_GUARD(spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx))
                   .set_stream_chunk_bytes(bytes))
}

SPLA_EXPORT SplaError spla_amd_ctx_set_gemm_mode(SplaContext ctx, SplaAmdGemmMode mode) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  if (mode != SPLA_AMD_GEMM_FP64 && mode != SPLA_AMD_GEMM_SPLIT_INT8 && mode != SPLA_AMD_GEMM_AUTO)
    return SPLA_INVALID_PARAMETER_ERROR;
  SPLA_C_GUARD(spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx))
                   .set_gemm_mode(static_cast<int>(mode)))
}

SPLA_EXPORT SplaError spla_amd_ctx_set_split_tolerance(SplaContext ctx, double tolerance) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  if (!(tolerance > 0.0)) return SPLA_INVALID_PARAMETER_ERROR;
  SPLA_C_GUARD(spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx))
                   .set_split_tolerance(tolerance))
}

SPLA_EXPORT SplaError spla_amd_ctx_set_split_workspace_limit(SplaContext ctx, size_t bytes) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx))
                   .set_split_workspace_limit(bytes))
}

SPLA_EXPORT SplaError spla_amd_ctx_last_product_info(SplaContext ctx, int* used, int* reason,
                                                     int* moduli, int* bitsA, int* bitsB,
                                                     double* bound) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  const spla_amd::SplitReport& r =
      spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx)).last_split();
  if (used) *used = r.used;
  if (reason) *reason = r.reason;
  if (moduli) *moduli = r.moduli;
  if (bitsA) *bitsA = r.bitsA;
  if (bitsB) *bitsB = r.bitsB;
  if (bound) *bound = r.bound;
  return SPLA_SUCCESS;
}

SPLA_EXPORT SplaError spla_amd_ctx_set_timing(SplaContext ctx, int enable) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx))
                   .timer()
                   .set_enabled(enable != 0))
}

SPLA_EXPORT SplaError spla_amd_ctx_last_timing(SplaContext ctx, double* ms) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  if (!ms) return SPLA_INVALID_POINTER_ERROR;
  const auto& t = spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx)).timer().ms();
  for (int i = 0; i < SPLA_AMD_NUM_PHASES; ++i) ms[i] = t[i];
  return SPLA_SUCCESS;
}

SPLA_EXPORT SplaError spla_amd_batch_create_ssb(SplaAmdBatch* batch, SplaContext ctx,
                                                SplaAmdElement type, SplaOperation opA, int count,
                                                const SplaAmdBlock* blocks, const void* alpha,
                                                const void* beta, SplaAmdTuning tuning) {
  if (!batch || !ctx) return SPLA_INVALID_HANDLE_ERROR;
  if (count < 0 || (type != SPLA_AMD_DOUBLE && type != SPLA_AMD_COMPLEX_DOUBLE) ||
      (tuning != SPLA_AMD_TUNE_MODEL && tuning != SPLA_AMD_TUNE_MEASURE &&
       tuning != SPLA_AMD_TUNE_PERSISTENT))
    return SPLA_INVALID_PARAMETER_ERROR;
  if ((count > 0 && !blocks) || !alpha || !beta) return SPLA_INVALID_POINTER_ERROR;
  SPLA_C_GUARD({
    std::vector<spla_amd::BlockSpec> specs(count);
    for (int i = 0; i < count; ++i) {
      const SplaAmdBlock& b = blocks[i];
      specs[i] = spla_amd::BlockSpec{b.m,   b.n,   b.k,   b.A,          b.lda,        b.B,
                                     b.ldb, b.C,   b.ldc, static_cast<int>(b.fillMode),
                                     b.rowOffset, b.colOffset};
    }
    auto ctxPtr = spla::InternalAccess::ctx_ptr(*reinterpret_cast<spla::Context*>(ctx));
    *batch = new spla_amd::BatchPlan(ctxPtr, type == SPLA_AMD_COMPLEX_DOUBLE, opA,
                                     std::move(specs), static_cast<const double*>(alpha),
                                     static_cast<const double*>(beta),
                                     static_cast<int>(tuning));
  })
}

SPLA_EXPORT SplaError spla_amd_batch_create_sbs(SplaAmdBatch* batch, SplaContext ctx,
                                                SplaAmdElement type, int count,
                                                const SplaAmdBlock* blocks, const void* alpha,
                                                const void* beta, SplaAmdTuning tuning) {
  return spla_amd_batch_create_ssb(batch, ctx, type, SPLA_OP_NONE, count, blocks, alpha, beta,
                                   tuning);
}

SPLA_EXPORT SplaError spla_amd_batch_persistent(SplaAmdBatch batch, int* persistent) {
  if (!batch) return SPLA_INVALID_HANDLE_ERROR;
  if (!persistent) return SPLA_INVALID_POINTER_ERROR;
  SPLA_C_GUARD(*persistent = static_cast<spla_amd::BatchPlan*>(batch)->persistent() ? 1 : 0)
}

SPLA_EXPORT SplaError spla_amd_batch_split_blocks(SplaAmdBatch batch, int* blocks) {
  if (!batch) return SPLA_INVALID_HANDLE_ERROR;
  if (!blocks) return SPLA_INVALID_POINTER_ERROR;
  *blocks = static_cast<const spla_amd::BatchPlan*>(batch)->last_split_blocks();
  return SPLA_SUCCESS;
}

SPLA_EXPORT SplaError spla_amd_batch_execute(SplaAmdBatch batch) {
  if (!batch) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(static_cast<spla_amd::BatchPlan*>(batch)->execute())
}

SPLA_EXPORT SplaError spla_amd_batch_info(SplaAmdBatch batch, int* tileConfig, int* items,
                                          int* workgroups, double* modelMs, double* measuredMs) {
  if (!batch) return SPLA_INVALID_HANDLE_ERROR;
  const auto* p = static_cast<const spla_amd::BatchPlan*>(batch);
  if (tileConfig) *tileConfig = p->tile_config();
  if (items) *items = p->num_items();
  if (workgroups) *workgroups = p->workgroups();
  if (modelMs) *modelMs = p->model_ms();
  if (measuredMs) *measuredMs = p->measured_ms();
  return SPLA_SUCCESS;
}

SPLA_EXPORT SplaError spla_amd_batch_destroy(SplaAmdBatch* batch) {
  if (!batch) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD({
    delete static_cast<spla_amd::BatchPlan*>(*batch);
    *batch = nullptr;
  })
}

SPLA_EXPORT SplaError spla_amd_ssb_ring_blocks(SplaDistributionType distType, int gridRows,
                                               int gridCols, int commSize, int m, int n,
                                               int cRowOffset, int cColOffset,
                                               SplaFillMode cFillMode, int tileSizeHost,
                                               int* rowsInBlock, int* colsInBlock,
                                               int* numBlocks, SplaAmdRingBlock* blocks,
                                               int capacity) {
  if ((distType != SPLA_DIST_BLACS_BLOCK_CYCLIC && distType != SPLA_DIST_MIRROR) ||
      gridRows < 1 || gridCols < 1 || commSize < 1 || m < 0 || n < 0 || cRowOffset < 0 ||
      cColOffset < 0 || tileSizeHost < 1 || capacity < 0 ||
      (cFillMode != SPLA_FILL_MODE_FULL && cFillMode != SPLA_FILL_MODE_UPPER &&
       cFillMode != SPLA_FILL_MODE_LOWER))
    return SPLA_INVALID_PARAMETER_ERROR;
  if (!numBlocks || (capacity > 0 && !blocks)) return SPLA_INVALID_POINTER_ERROR;
  SPLA_C_GUARD({
    const auto p = spla_amd::plan_ssb_ring(distType == SPLA_DIST_MIRROR, gridRows, gridCols,
                                           commSize, m, n, cRowOffset, cColOffset, cFillMode,
                                           tileSizeHost);
    if (rowsInBlock) *rowsInBlock = static_cast<int>(p.rowsInBlock);
    if (colsInBlock) *colsInBlock = static_cast<int>(p.colsInBlock);
    *numBlocks = static_cast<int>(p.blocks.size());
    const std::size_t w = std::min<std::size_t>(p.blocks.size(), static_cast<std::size_t>(capacity));
    for (std::size_t i = 0; i < w; ++i) {
      const auto& b = p.blocks[i];
      blocks[i] = SplaAmdRingBlock{static_cast<int>(b.row), static_cast<int>(b.col),
                                   static_cast<int>(b.rows), static_cast<int>(b.cols)};
    }
  })
}

SPLA_EXPORT SplaError spla_amd_region_pieces(SplaDistributionType distType, int rowsInBlock,
                                             int colsInBlock, int gridRows, int gridCols,
                                             int numRows, int numCols, int rowOffset,
                                             int colOffset, int* numPieces, SplaAmdPiece* pieces,
                                             int capacity) {
  if ((distType != SPLA_DIST_BLACS_BLOCK_CYCLIC && distType != SPLA_DIST_MIRROR) ||
      rowsInBlock < 1 || colsInBlock < 1 || gridRows < 1 || gridCols < 1 || numRows < 0 ||
      numCols < 0 || rowOffset < 0 || colOffset < 0 || capacity < 0)
    return SPLA_INVALID_PARAMETER_ERROR;
  if (!numPieces || (capacity > 0 && !pieces)) return SPLA_INVALID_POINTER_ERROR;
  SPLA_C_GUARD({
    const auto v = spla_amd::region_pieces(distType == SPLA_DIST_MIRROR, rowsInBlock, colsInBlock,
                                           gridRows, gridCols, numRows, numCols, rowOffset,
                                           colOffset);
    *numPieces = static_cast<int>(v.size());
    const std::size_t w = std::min<std::size_t>(v.size(), static_cast<std::size_t>(capacity));
    for (std::size_t i = 0; i < w; ++i) {
      const auto& q = v[i];
      pieces[i] = SplaAmdPiece{static_cast<int>(q.globalRow), static_cast<int>(q.globalCol),
                               static_cast<int>(q.subRow),    static_cast<int>(q.subCol),
                               static_cast<int>(q.localRow),  static_cast<int>(q.localCol),
                               static_cast<int>(q.rows),      static_cast<int>(q.cols),
                               static_cast<int>(q.rank)};
    }
  })
}

}  // extern "C"
