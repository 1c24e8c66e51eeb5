// Element-wise helper kernels of the engine: precision conversion of 2-D
// column-major blocks (single-precision entry points run on the FP64 kernels).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels/launch.hpp"

namespace spla_amd {
namespace {

template <typename S, typename D>
__global__ void __launch_bounds__(256) convert_kernel(const S* __restrict__ src, int64_t srcLd,
                                                      D* __restrict__ dst, int64_t dstLd,
                                                      int64_t rows, int64_t cols) {
  const int64_t total = rows * cols;
  for (int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; idx < total;
       idx += static_cast<int64_t>(gridDim.x) * 256) {
    const int64_t r = idx % rows;
    const int64_t c = idx / rows;
    dst[c * dstLd + r] = static_cast<D>(src[c * srcLd + r]);
  }
}

template <typename S, typename D>
hipError_t launch_convert(const S* src, int64_t srcLd, D* dst, int64_t dstLd, int64_t rows,
                          int64_t cols, hipStream_t stream) {
  const int64_t total = rows * cols;
  if (total == 0) return hipSuccess;
  const int64_t blocks = (total + 255) / 256;
  const unsigned grid = static_cast<unsigned>(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL((convert_kernel<S, D>), dim3(grid), dim3(256), 0, stream, src, srcLd, dst,
                     dstLd, rows, cols);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_convert_f32_f64(const float* src, int64_t srcLd, double* dst, int64_t dstLd,
                                  int64_t rows, int64_t cols, hipStream_t stream) {
  return launch_convert(src, srcLd, dst, dstLd, rows, cols, stream);
}

hipError_t launch_convert_f64_f32(const double* src, int64_t srcLd, float* dst, int64_t dstLd,
                                  int64_t rows, int64_t cols, hipStream_t stream) {
  return launch_convert(src, srcLd, dst, dstLd, rows, cols, stream);
}

}  // namespace spla_amd
