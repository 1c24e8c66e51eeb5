// Host entry points of the HIP kernels. Every function enqueues work on the
// given stream and returns the launch status; none of them synchronizes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels/batched_desc.hpp"
#include "kernels/gemm_desc.hpp"
#include "kernels/ozaki_desc.hpp"

namespace spla_amd {

// FP64 / complex-FP64 MFMA GEMM over the tiles described by d (gemm_f64.hip).
hipError_t launch_gemm_f64(const GemmDesc& d, bool complex, int layoutA, int layoutB, int cfg,
                           hipStream_t stream);
// The same in FP32 / complex-FP32 (FP32 accumulation; d's pointers are floats).
hipError_t launch_gemm_f32(const GemmDesc& d, bool complex, int layoutA, int layoutB, int cfg,
                           hipStream_t stream);

// Streaming FP64 / complex-FP64 product for small m, n with K-major A and B
// (skinny.hip): one workgroup per split (gridDim = d.splits >= 2), each summing
// its d.kSplit embedded reduction indices into d.work ([splits][n][Mp]) for
// launch_splitk_reduce.
bool skinny_supported(bool complex, int layoutA, int layoutB, int64_t m, int64_t n);
hipError_t launch_skinny_gemm(const GemmDesc& d, bool complex, hipStream_t stream);

// Fixed-order sum of d.splits partial results in d.work, then alpha / beta
// (f32: the scalars of d are floats).
hipError_t launch_splitk_reduce(const GemmDesc& d, bool complex, int cfg, hipStream_t stream,
                                bool f32 = false);

// C <- beta * C (beta == 0: C <- 0 without reading C).
hipError_t launch_scale(const GemmDesc& d, bool complex, hipStream_t stream, bool f32 = false);

// Split-precision kernels (ozaki.hip).
// Per-row largest |component| and sum of squares of an operand read as `rows`
// rows of k elements; `partial` holds 2 * splits * rows doubles.
hipError_t launch_split_stats(bool complex, int layout, const double* X, int64_t ld, int rows,
                              int64_t k, double* partial, int splits, double* outMax,
                              double* outSum, hipStream_t stream);
hipError_t launch_residues(bool complex, int layout, const ResidueDesc& d, hipStream_t stream);
hipError_t launch_i8_gemm(const I8GemmDesc& d, int cfg, hipStream_t stream);
// limbs: CrtConstants::limbs of the constants d.cc points to
hipError_t launch_crt(bool complex, const CrtDesc& d, int limbs, hipStream_t stream);

// Batched block products (batched.hip): the GEMM kernel (grouped: one workgroup
// per BatchItem; persistent: one workgroup per slot of the static schedule;
// kernel argument one BatchDesc) and the split reduction kernel (one workgroup
// per BatchReduceTile) of a tile configuration (TileL or TileM) and A layout
// (KMajor: ssb blocks, MNMajor: sbs blocks), as kernel handles for
// hipLaunchKernel and hipGraphAddKernelNode; `threads` per workgroup.
struct BatchKernels {
  const void* gemm;
  const void* reduce;
  int threads;
};
BatchKernels batched_kernels(bool complex, int layoutA, int cfg, bool persistent);

}  // namespace spla_amd
