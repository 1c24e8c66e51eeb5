// spla::Context and the spla_ctx_* C entry points
// (SPLA src/spla/context.cpp:39-366).
#include <memory>

#include "api/internal_access.hpp"
#include "spla/context.h"
#include "spla/context.hpp"

namespace spla {

Context::Context(SplaProcessingUnit pu) : ctxInternal_(std::make_shared<ContextInternal>(pu)) {}

SplaProcessingUnit Context::processing_unit() const { return ctxInternal_->processing_unit(); }
int Context::num_tiles() const { return ctxInternal_->num_tiles(); }
int Context::tile_size_host() const { return ctxInternal_->tile_size_host(); }
int Context::tile_size_gpu() const { return ctxInternal_->tile_size_gpu(); }
int Context::op_threshold_gpu() const { return ctxInternal_->op_threshold_gpu(); }
int Context::gpu_device_id() const { return ctxInternal_->gpu_device_id(); }

std::uint_least64_t Context::allocated_memory_host() const {
  return ctxInternal_->host_memory()->size();
}
std::uint_least64_t Context::allocated_memory_pinned() const {
  return ctxInternal_->pinned_memory()->size();
}
std::uint_least64_t Context::allocated_memory_gpu() const {
  return ctxInternal_->gpu_memory()->size();
}

void Context::set_num_tiles(int numTiles) { ctxInternal_->set_num_tiles(numTiles); }
void Context::set_tile_size_host(int v) { ctxInternal_->set_tile_size_host(v); }
void Context::set_tile_size_gpu(int v) { ctxInternal_->set_tile_size_gpu(v); }
void Context::set_op_threshold_gpu(int v) { ctxInternal_->set_op_threshold_gpu(v); }

namespace {
std::shared_ptr<spla_amd::MemoryResource> user_resource(
    std::function<void*(std::size_t)>& a, std::function<void(void*)>& f) {
  if (!a || !f) throw InvalidAllocatorFunctionError();
  return std::make_shared<spla_amd::UserResource>(std::move(a), std::move(f));
}
}  // namespace

void Context::set_alloc_host(std::function<void*(std::size_t)> a, std::function<void(void*)> f) {
  ctxInternal_->set_host_memory(user_resource(a, f));
}
void Context::set_alloc_pinned(std::function<void*(std::size_t)> a,
                               std::function<void(void*)> f) {
  ctxInternal_->set_pinned_memory(user_resource(a, f));
}
void Context::set_alloc_gpu(std::function<void*(std::size_t)> a, std::function<void(void*)> f) {
  ctxInternal_->set_gpu_memory(user_resource(a, f));
}

}  // namespace spla

namespace {
spla::Context& as_ctx(SplaContext ctx) { return *reinterpret_cast<spla::Context*>(ctx); }
}  // namespace

extern "C" {

SPLA_EXPORT SplaError spla_ctx_create(SplaContext* ctx, SplaProcessingUnit pu) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(*ctx = reinterpret_cast<void*>(new spla::Context(pu)))
}

SPLA_EXPORT SplaError spla_ctx_destroy(SplaContext* ctx) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  try {
    delete reinterpret_cast<spla::Context*>(*ctx);
  } catch (const spla::GenericError& e) {
    return e.error_code();
  } catch (...) {
    return SPLA_UNKNOWN_ERROR;
  }
  *ctx = nullptr;
  return SPLA_SUCCESS;
}

#define SPLA_CTX_GETTER(NAME, TYPE, EXPR)                 \
  SPLA_EXPORT SplaError NAME(SplaContext ctx, TYPE* out) { \
    if (!ctx) return SPLA_INVALID_HANDLE_ERROR;            \
    SPLA_C_GUARD(*out = as_ctx(ctx).EXPR)                  \
  }

SPLA_CTX_GETTER(spla_ctx_processing_unit, SplaProcessingUnit, processing_unit())
SPLA_CTX_GETTER(spla_ctx_num_tiles, int, num_tiles())
SPLA_CTX_GETTER(spla_ctx_tile_size_host, int, tile_size_host())
SPLA_CTX_GETTER(spla_ctx_tile_size_gpu, int, tile_size_gpu())
SPLA_CTX_GETTER(spla_ctx_op_threshold_gpu, int, op_threshold_gpu())
SPLA_CTX_GETTER(spla_ctx_gpu_device_id, int, gpu_device_id())
SPLA_CTX_GETTER(spla_ctx_allocated_memory_host, uint_least64_t, allocated_memory_host())
SPLA_CTX_GETTER(spla_ctx_allocated_memory_pinned, uint_least64_t, allocated_memory_pinned())
SPLA_CTX_GETTER(spla_ctx_allocated_memory_gpu, uint_least64_t, allocated_memory_gpu())
#undef SPLA_CTX_GETTER

#define SPLA_CTX_SETTER(NAME, METHOD)                   \
  SPLA_EXPORT SplaError NAME(SplaContext ctx, int value) { \
    if (!ctx) return SPLA_INVALID_HANDLE_ERROR;          \
    SPLA_C_GUARD(as_ctx(ctx).METHOD(value))              \
  }

SPLA_CTX_SETTER(spla_ctx_set_num_tiles, set_num_tiles)
SPLA_CTX_SETTER(spla_ctx_set_tile_size_host, set_tile_size_host)
SPLA_CTX_SETTER(spla_ctx_set_op_threshold_gpu, set_op_threshold_gpu)
SPLA_CTX_SETTER(spla_ctx_set_tile_size_gpu, set_tile_size_gpu)
#undef SPLA_CTX_SETTER

#define SPLA_CTX_ALLOC(NAME, METHOD)                                                      \
  SPLA_EXPORT SplaError NAME(SplaContext ctx, void* (*allocateFunc)(size_t),             \
                             void (*deallocateFunc)(void*)) {                            \
    if (!ctx) return SPLA_INVALID_HANDLE_ERROR;                                          \
    if (!allocateFunc || !deallocateFunc) return SPLA_INVALID_ALLOCATOR_FUNCTION;        \
    SPLA_C_GUARD(as_ctx(ctx).METHOD(allocateFunc, deallocateFunc))                       \
  }

SPLA_CTX_ALLOC(spla_ctx_set_alloc_host, set_alloc_host)
SPLA_CTX_ALLOC(spla_ctx_set_alloc_pinned, set_alloc_pinned)
SPLA_CTX_ALLOC(spla_ctx_set_alloc_gpu, set_alloc_gpu)
#undef SPLA_CTX_ALLOC

}  // extern "C"
