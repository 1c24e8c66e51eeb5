// Local GEMM engine: the MI355X counterpart of SPLA's gemm_gpu
// (src/gemm/gemm_gpu.cpp:59-174) and multiply_gpu
// (src/gpu_util/multiply_gpu.cpp:78-143).
//
//   * parameters are validated as SPLA's check_gemm_param
//     (src/util/check_gemm_param.hpp:36-50);
//   * the call is synchronous: the legacy default stream is synchronized on
//     entry (gemm_gpu.cpp:72) and the engine's streams before returning;
//   * operands are classified with hipPointerGetAttributes
//     (gpu_pointer_translation.hpp:41-81): device memory is used in place,
//     anything else is host memory;
//   * host-resident A / B are streamed in k-chunks on a copy stream into
//     double-buffered device chunks; pageable memory is packed by host threads
//     into pinned staging buffers first, pinned memory is copied directly.
//     Copies of chunk c+1 overlap the MFMA kernels of chunk c. Host-resident C
//     is copied in (if beta != 0) and out once;
//   * every product runs on the HIP MFMA kernels; there is no host compute path.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

#include "runtime/allocators.hpp"
#include "runtime/context_internal.hpp"
#include "spla/types.h"

namespace spla_amd {

// Part of C that has to be computed (pgemm_ssbtr). Tiles of C wholly outside
// the triangle are skipped; rowOff / colOff locate C in the global matrix.
struct FillSpec {
  int mode = SPLA_FILL_MODE_FULL;
  int64_t rowOff = 0, colOff = 0;
};

// T: float, double, std::complex<float>, std::complex<double>.
template <typename T>
void gemm_gpu(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k, T alpha,
              const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc,
              spla::ContextInternal& ctx, FillSpec fill = FillSpec());

// One product on device-resident operands on the FP64 kernels with the plan
// gemm_gpu makes for it in FP64 mode, its split-K partial sums in the caller's
// `workspace` (256-byte aligned): no device memory is allocated. Throws
// GPUInvalidDevicePointerError for host operands and GPUAllocationError when
// the plan's partial sums need more than workspaceBytes. Synchronous like
// gemm_gpu (T: std::complex<double>).
template <typename T>
void gemm_gpu_fixed_workspace(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n,
                              int64_t k, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                              T beta, T* C, int64_t ldc, spla::ContextInternal& ctx,
                              void* workspace, std::size_t workspaceBytes);

// One product on device-resident operands, enqueued on `stream` with no
// synchronization (T: double or std::complex<double>). Workspace blocks are
// appended to `keep` and must live until the stream has drained; `reuse`
// (optional) is one split-K workspace shared by products on the same stream.
// trySplit = false runs the FP64 kernels whatever the context's mode;
// wholeChip = false plans for a device shared with other streams (the block
// products of a batch).
template <typename T>
void enqueue_on_device(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                       T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C,
                       int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill,
                       hipStream_t stream, std::vector<Block>& keep, Block* reuse,
                       bool trySplit = true, bool wholeChip = false);

// Estimated seconds of the FP64 kernels for one product (the tuner's model),
// which the automatic mode compares with the split form; wholeChip = false for
// the block products of a batch, which share the device.
double fp64_estimate_seconds(spla::ContextInternal& ctx, bool complex, SplaOperation opA,
                             SplaOperation opB, int64_t m, int64_t n, int64_t k,
                             bool wholeChip = true);

// SPLA's check_gemm_param (src/util/check_gemm_param.hpp:36-50): throws
// InvalidParameterError / InvalidPointerError.
void check_gemm_param(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                      const void* A, int64_t lda, const void* B, int64_t ldb, const void* C,
                      int64_t ldc);

}  // namespace spla_amd
