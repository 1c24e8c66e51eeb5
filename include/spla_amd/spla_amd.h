/*
 * Extensions of this library beyond SPLA's C API. None of these entry points
 * has a counterpart in SPLA; they are for callers that integrate the engine
 * with their own HIP streams or select the arithmetic of the products.
 */
#ifndef SPLA_AMD_H
#define SPLA_AMD_H

#include "spla/config.h"
#include "spla/context.h"
#include "spla/errors.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Run the context's kernels on the caller's HIP stream (a hipStream_t passed
 * as void*; NULL restores the context's own stream). Calls stay synchronous:
 * they return after the stream has drained. Intended for callers that time or
 * order the library's kernels with their own events (e.g. a torch stream).
 */
SPLA_EXPORT SplaError spla_amd_ctx_set_stream(SplaContext ctx, void* hipStream);

/*
 * Target size in bytes of one k-chunk of host-resident operands streamed to
 * the GPU (default 256 MiB; at least 64 reduction indices per chunk).
 * Smaller chunks start the pipeline earlier, larger ones run longer kernels.
 */
SPLA_EXPORT SplaError spla_amd_ctx_set_stream_chunk_bytes(SplaContext ctx, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif
