// Batched FP64 / complex-FP64 block products of one ring step: one GEMM launch
// over the (block, tile, split) items of every block, then one reduction launch
// over the tiles whose reduction is split, which adds the partial sums in split
// order (bitwise reproducible) and applies alpha / beta. The GEMM launch is
// grouped (one workgroup per item, in decreasing cost order: the dispatcher
// hands them to the CUs as these free up) or persistent (one workgroup per
// resident slot, each working through the static list the host's
// longest-processing-time schedule gave it). Both launches read only
// descriptors and operands; the partial sums are written before they are read
// in every launch, so the pair can be replayed from a hipGraph without clearing
// anything.
//
// ssb blocks (SPLA src/pgemm_ssb/ring_ssb_gpu.cpp:193-197): C_b <- alpha op(A_b) B_b
// + beta C_b with op(A) = A^T / A^H, A k x m (K-major). sbs blocks
// (src/pgemm_sbs/ring_sbs_gpu.cpp:57-88): op(A) = A, A m x k (MN-major).
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

// One (block, tile, split) item; on return every thread has passed the tile
// routine's last barrier, so the LDS buffers may be reused at once.
template <bool CPLX, int LA, int WM, int WN, bool LEAN>
__device__ __forceinline__ void batched_item(
    const BatchDesc& d, const BatchItem& it,
    typename TileMma<double, CPLX, LA, KMajor, WM, WN, 2, 16>::Smem& sm) {
  using T = TileMma<double, CPLX, LA, KMajor, WM, WN, 2, 16, LEAN>;
  const BatchBlock b = d.blocks[it.block];
  const int row0 = it.tm * T::BM;  // embedded rows
  const int col0 = it.tn * T::BN;
  const int64_t kBegin = static_cast<int64_t>(it.s) * b.kSplit;
  const int64_t kEnd = min(b.Kp, kBegin + b.kSplit);
  const TileOperands<double> op{b.A, b.B, b.lda, b.ldb, b.m, b.n, d.conjA, d.conjB};
  dbl4 acc[T::FM][T::FN];
  T::run(op, row0, col0, kBegin, kEnd, sm, acc);
  if (b.splits == 1) {
    T::for_each(row0, col0, b.m, b.n, acc,
                [&](int i, int j, double re, double im) { update_c<CPLX>(d, b, i, j, re, im); });
  } else {
    double* w = d.work + b.workOffset + static_cast<int64_t>(it.s) * b.n * b.Mp;
    T::for_each(row0, col0, b.m, b.n, acc, [&](int i, int j, double re, double im) {
      store_work<double, CPLX>(w, b.Mp, i, j, re, im);
    });
  }
}

// Grouped launch: workgroup = item (gridDim.x = number of items).
template <bool CPLX, int LA, int WM, int WN>
__global__ void __launch_bounds__(kThreads, 2) batched_gemm_kernel(BatchDesc d) {
  __shared__ typename TileMma<double, CPLX, LA, KMajor, WM, WN, 2, 16>::Smem sm;
  batched_item<CPLX, LA, WM, WN, false>(d, d.items[blockIdx.x], sm);
}

// Persistent launch: one workgroup per slot works through its static list of
// items (gridDim.x = number of slots); no workgroup waits for another.
template <bool CPLX, int LA, int WM, int WN>
__global__ void __launch_bounds__(kThreads, 2) batched_gemm_persistent_kernel(BatchDesc d) {
  __shared__ typename TileMma<double, CPLX, LA, KMajor, WM, WN, 2, 16>::Smem sm;
  const int end = d.slotStart[blockIdx.x + 1];
  for (int i = d.slotStart[blockIdx.x]; i < end; ++i)
    batched_item<CPLX, LA, WM, WN, true>(d, d.items[d.slotItems[i]], sm);
}

// AR x BN: element rows and columns of a tile of the GEMM launch.
template <bool CPLX, int AR, int BN>
__global__ void __launch_bounds__(kThreads) batched_reduce_kernel(BatchDesc d) {
  const BatchReduceTile t = d.reduce[blockIdx.x];
  const BatchBlock b = d.blocks[t.block];
  const int i0 = t.tm * AR;
  const int j0 = t.tn * BN;
  const int rows = min(AR, b.m - i0);
  const int cols = min(BN, b.n - j0);
  const int64_t stride = static_cast<int64_t>(b.n) * b.Mp;
  const double* w0 = d.work + b.workOffset;
  for (int e = threadIdx.x; e < rows * cols; e += kThreads) {
    const int i = i0 + e % rows;
    const int j = j0 + e / rows;
    const double* p = w0 + static_cast<int64_t>(j) * b.Mp + (CPLX ? 2 * i : i);
    double re = 0.0, im = 0.0;
    for (int s = 0; s < b.splits; ++s) {
      re += p[s * stride];
      if (CPLX) im += p[s * stride + 1];
    }
    update_c<CPLX>(d, b, i, j, re, im);
  }
}

template <bool CPLX, int LA>
const void* gemm_kernel_for(int cfg, bool persistent) {
  if (persistent) {
    if (cfg == TileM)
      return reinterpret_cast<const void*>(&batched_gemm_persistent_kernel<CPLX, LA, 32, 32>);
    return reinterpret_cast<const void*>(&batched_gemm_persistent_kernel<CPLX, LA, 64, 64>);
  }
  if (cfg == TileM) return reinterpret_cast<const void*>(&batched_gemm_kernel<CPLX, LA, 32, 32>);
  return reinterpret_cast<const void*>(&batched_gemm_kernel<CPLX, LA, 64, 64>);
}

}  // namespace

BatchKernels batched_kernels(bool complex, int layoutA, int cfg, bool persistent) {
  BatchKernels k;
  if (complex) {
    k.gemm = layoutA == MNMajor ? gemm_kernel_for<true, MNMajor>(cfg, persistent)
                                : gemm_kernel_for<true, KMajor>(cfg, persistent);
    k.reduce = cfg == TileM ? reinterpret_cast<const void*>(&batched_reduce_kernel<true, 32, 64>)
                            : reinterpret_cast<const void*>(&batched_reduce_kernel<true, 64, 128>);
  } else {
    k.gemm = layoutA == MNMajor ? gemm_kernel_for<false, MNMajor>(cfg, persistent)
                                : gemm_kernel_for<false, KMajor>(cfg, persistent);
    k.reduce = cfg == TileM ? reinterpret_cast<const void*>(&batched_reduce_kernel<false, 64, 64>)
                            : reinterpret_cast<const void*>(&batched_reduce_kernel<false, 128, 128>);
  }
  k.threads = kThreads;
  return k;
}

}  // namespace spla_amd
