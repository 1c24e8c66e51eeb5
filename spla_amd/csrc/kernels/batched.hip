// Persistent batched FP64 / complex-FP64 block products for one ring step:
// every workgroup pulls (block, tile, split) items from a device queue until it
// is empty, so hundreds of uneven products share one launch (and one hipGraph
// node). Items are ordered by decreasing cost on the host.
//
// Split tiles: each split writes its raw partial sums to the block's
// workspace, then increments the tile's arrival counter; the workgroup that
// arrives last adds the partials in split order (bitwise reproducible) and
// applies alpha / beta. Counters and the queue are returned to zero by the
// last user, so the launch can be replayed without memset nodes.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels/batched_desc.hpp"
#include "kernels/launch.hpp"
#include "kernels/tile_mma.hpp"

namespace spla_amd {
namespace {

using namespace tile;

template <bool CPLX>
__device__ __forceinline__ void update_c(const BatchDesc& d, const BatchBlock& b, int i, int j,
                                         double re, double im) {
  if (CPLX) {
    double* c = b.C + (static_cast<int64_t>(j) * b.ldc + i) * 2;
    double outRe = d.alphaRe * re - d.alphaIm * im;
    double outIm = d.alphaRe * im + d.alphaIm * re;
    if (!d.betaZero) {
      const double cr = c[0], ci = c[1];
      outRe += d.betaRe * cr - d.betaIm * ci;
      outIm += d.betaRe * ci + d.betaIm * cr;
    }
    c[0] = outRe;
    c[1] = outIm;
  } else {
    double* c = b.C + static_cast<int64_t>(j) * b.ldc + i;
    double out = d.alphaRe * re;
    if (!d.betaZero) out += d.betaRe * c[0];
    c[0] = out;
  }
}

template <bool CPLX, int WM, int WN>
__global__ void __launch_bounds__(kThreads, 2) batched_ssb_kernel(BatchDesc d) {
  using T = TileMma<double, CPLX, KMajor, KMajor, WM, WN, 2, 16>;
  __shared__ typename T::Smem sm;
  __shared__ int sItem;
  __shared__ int sLast;

  for (;;) {
    if (threadIdx.x == 0) sItem = static_cast<int>(atomicAdd(&d.queue[0], 1u));
    __syncthreads();
    const int idx = sItem;
    __syncthreads();
    if (idx >= d.numItems) break;
    const BatchItem it = d.items[idx];
    const BatchBlock b = d.blocks[it.block];
    const int row0 = it.tm * T::BM;
    const int col0 = it.tn * T::BN;
    const int64_t kBegin = static_cast<int64_t>(it.s) * b.kSplit;
    const int64_t kEnd = min(b.Kp, kBegin + b.kSplit);
    const TileOperands<double> op{b.A, b.B, b.lda, b.ldb, b.m, b.n, d.conjA, d.conjB};
    dbl4 acc[T::FM][T::FN];
    T::run(op, row0, col0, kBegin, kEnd, sm, acc);

    if (b.splits == 1) {
      T::for_each(row0, col0, b.m, b.n, acc,
                  [&](int i, int j, double re, double im) { update_c<CPLX>(d, b, i, j, re, im); });
      continue;
    }
    double* w = d.work + b.workOffset + static_cast<int64_t>(it.s) * b.n * b.Mp;
    T::for_each(row0, col0, b.m, b.n, acc, [&](int i, int j, double re, double im) {
      store_work<double, CPLX>(w, b.Mp, i, j, re, im);
    });
    __threadfence();  // partial sums visible device-wide before the arrival
    __syncthreads();
    unsigned* counter = &d.counters[b.counterBase + it.tm * b.tilesN + it.tn];
    if (threadIdx.x == 0) sLast = atomicAdd(counter, 1u) == static_cast<unsigned>(b.splits - 1);
    __syncthreads();
    if (!sLast) continue;
    __threadfence();  // acquire: observe every split's partial sums
    const int rows = min(T::AR, b.m - (CPLX ? row0 / 2 : row0));
    const int cols = min(T::BN, b.n - col0);
    const int i0 = CPLX ? row0 / 2 : row0;
    const int64_t stride = static_cast<int64_t>(b.n) * b.Mp;
    const double* w0 = d.work + b.workOffset;
    for (int e = threadIdx.x; e < rows * cols; e += kThreads) {
      const int i = i0 + e % rows;
      const int j = col0 + e / rows;
      const double* p = w0 + static_cast<int64_t>(j) * b.Mp + (CPLX ? 2 * i : i);
      double re = 0.0, im = 0.0;
      for (int s = 0; s < b.splits; ++s) {
        re += p[s * stride];
        if (CPLX) im += p[s * stride + 1];
      }
      update_c<CPLX>(d, b, i, j, re, im);
    }
    if (threadIdx.x == 0) atomicExch(counter, 0u);  // ready for the next launch
  }
  // The last workgroup to leave resets the queue for the next launch.
  if (threadIdx.x == 0) {
    if (atomicAdd(&d.queue[1], 1u) == gridDim.x - 1) {
      atomicExch(&d.queue[0], 0u);
      atomicExch(&d.queue[1], 0u);
    }
  }
}

}  // namespace

hipError_t launch_batched_ssb(const BatchDesc& d, bool complex, int cfg, int workgroups,
                              hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>(workgroups)), block(kThreads);
  if (complex) {
    if (cfg == TileM)
      hipLaunchKernelGGL((batched_ssb_kernel<true, 32, 32>), grid, block, 0, stream, d);
    else
      hipLaunchKernelGGL((batched_ssb_kernel<true, 64, 64>), grid, block, 0, stream, d);
  } else {
    if (cfg == TileM)
      hipLaunchKernelGGL((batched_ssb_kernel<false, 32, 32>), grid, block, 0, stream, d);
    else
      hipLaunchKernelGGL((batched_ssb_kernel<false, 64, 64>), grid, block, 0, stream, d);
  }
  return hipGetLastError();
}

}  // namespace spla_amd
