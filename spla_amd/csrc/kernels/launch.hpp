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

// Persistent batched block products (batched.hip); cfg is TileL or TileM.
hipError_t launch_batched_ssb(const BatchDesc& d, bool complex, int cfg, int workgroups,
                              hipStream_t stream);

}  // namespace spla_amd
