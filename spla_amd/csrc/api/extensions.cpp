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

}  // extern "C"
