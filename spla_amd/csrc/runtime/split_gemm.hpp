// Split-precision (Ozaki scheme II) products on the INT8 matrix cores.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <vector>

#include "runtime/allocators.hpp"
#include "runtime/context_internal.hpp"
#include "spla/types.h"

namespace spla_amd {

struct FillSpec;

// Enqueues C <- alpha op(A) op(B) + beta C for device-resident operands on the
// split-precision path and returns true, or returns false (nothing enqueued)
// when the plan falls back to the FP64 kernels. The decision needs the operand
// statistics on the host, so the stream is synchronized once inside. Work
// buffers are appended to `keep`.
template <typename T>
bool enqueue_split_product(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                           T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                           T* C, int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill,
                           double fp64Seconds, hipStream_t stream, std::vector<Block>& keep);

}  // namespace spla_amd
