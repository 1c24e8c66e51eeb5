// Entry points of include/spla_amd/spla_amd.h (no SPLA counterpart).
#include "api/internal_access.hpp"
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

}  // extern "C"
