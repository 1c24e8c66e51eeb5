/*
 * Context: configuration and reusable resources (streams, pools, tuning cache).
 *
 * C entry points of SPLA 1.6.1 include/spla/context.h:46-207, implemented
 * there in src/spla/context.cpp:122-365. Each function here replaces the
 * function of the same name.
 */
#ifndef SPLA_CONTEXT_H
#define SPLA_CONTEXT_H

#include <stddef.h>
#include <stdint.h>

#include "spla/config.h"
#include "spla/errors.h"
#include "spla/types.h"

/* Opaque handle holding a heap-allocated spla::Context. */
typedef void* SplaContext;

#ifdef __cplusplus
extern "C" {
#endif

/* Create a context on processing unit pu (context.cpp:122-131). */
SPLA_EXPORT SplaError spla_ctx_create(SplaContext* ctx, SplaProcessingUnit pu);

/* Destroy a context and set *ctx to NULL; NULL ctx -> SPLA_INVALID_HANDLE_ERROR (context.cpp:133-146). */
SPLA_EXPORT SplaError spla_ctx_destroy(SplaContext* ctx);

/* Processing unit of the context (declared at context.h:70 in SPLA, never defined there). */
SPLA_EXPORT SplaError spla_ctx_processing_unit(SplaContext ctx, SplaProcessingUnit* pu);

/* Number of concurrent tiles / streams used on the GPU (context.cpp:148-160). */
SPLA_EXPORT SplaError spla_ctx_num_tiles(SplaContext ctx, int* numTiles);

/* Target block size of distributed host work (context.cpp:162-174). */
SPLA_EXPORT SplaError spla_ctx_tile_size_host(SplaContext ctx, int* tileSizeHost);

/* Side length of the GPU staging tile for host-resident operands (context.cpp:176-188). */
SPLA_EXPORT SplaError spla_ctx_tile_size_gpu(SplaContext ctx, int* tileSizeGPU);

/* Operation-count threshold for small host-resident products (context.cpp:190-202). */
SPLA_EXPORT SplaError spla_ctx_op_threshold_gpu(SplaContext ctx, int* opThresholdGPU);

/* GPU device the context was created on (context.cpp:204-216). */
SPLA_EXPORT SplaError spla_ctx_gpu_device_id(SplaContext ctx, int* deviceId);

/* Bytes currently held by the host / pinned / GPU pools (context.cpp:218-258). */
SPLA_EXPORT SplaError spla_ctx_allocated_memory_host(SplaContext ctx, uint_least64_t* size);
SPLA_EXPORT SplaError spla_ctx_allocated_memory_pinned(SplaContext ctx, uint_least64_t* size);
SPLA_EXPORT SplaError spla_ctx_allocated_memory_gpu(SplaContext ctx, uint_least64_t* size);

/* Setters; values below 1 (below 0 for the threshold) -> SPLA_INVALID_PARAMETER_ERROR
 * (context.cpp:260-314, context_internal.hpp:122-140). */
SPLA_EXPORT SplaError spla_ctx_set_num_tiles(SplaContext ctx, int numTiles);
SPLA_EXPORT SplaError spla_ctx_set_tile_size_host(SplaContext ctx, int tileSizeHost);
SPLA_EXPORT SplaError spla_ctx_set_op_threshold_gpu(SplaContext ctx, int opThresholdGPU);
SPLA_EXPORT SplaError spla_ctx_set_tile_size_gpu(SplaContext ctx, int tileSizeGPU);

/* Replace the host / pinned / GPU pool with user functions; NULL functions ->
 * SPLA_INVALID_ALLOCATOR_FUNCTION (context.cpp:316-366). */
SPLA_EXPORT SplaError spla_ctx_set_alloc_host(SplaContext ctx, void* (*allocateFunc)(size_t),
                                              void (*deallocateFunc)(void*));
SPLA_EXPORT SplaError spla_ctx_set_alloc_pinned(SplaContext ctx, void* (*allocateFunc)(size_t),
                                                void (*deallocateFunc)(void*));
SPLA_EXPORT SplaError spla_ctx_set_alloc_gpu(SplaContext ctx, void* (*allocateFunc)(size_t),
                                             void (*deallocateFunc)(void*));

#ifdef __cplusplus
}
#endif

#endif
