// FP64 / complex-FP64 GEMM on the gfx950 FP64 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Replaces the vendor ?gemm call of SPLA's local engine
// (src/gpu_util/multiply_gpu.cpp:40-76 -> rocblas_{d,z}gemm) for the
// operations the engine issues: opA in {N, T, C}, opB in {N, T, C}.
//
// Structure of one workgroup (256 threads = 4 wave64):
//   * a BM x BN tile of the embedded output (see gemm_desc.hpp), split over a
//     WAVES_M x WAVES_N grid of waves, each owning WM x WN = (WM/16)x(WN/16)
//     MFMA accumulators of 16x16 doubles;
//   * the reduction runs over k-blocks of BK embedded elements, double
//     buffered in LDS with one barrier per k-block; the next block's global
//     loads are issued before the MFMAs of the current one;
//   * inside a k-block the reduction index is permuted so that every lane
//     reads its A and B operands as 16-byte pairs (q = 8*b + 2*kk + step):
//     A and B use the same permutation, so the sum is unchanged;
//   * complex: inside each 16-row MFMA fragment, rows 0-7 carry the real and
//     rows 8-15 the imaginary parts of 8 complex rows of C. The A operand of
//     an imaginary row reads the partner component of the pair (swap) with a
//     lane-constant sign, so A and B are staged exactly once and no operand
//     is expanded in memory; every lane ends up holding whole complex results;
//   * split-K: a split writes its raw partial sums to a workspace; the
//     reduction kernel adds the splits in a fixed order (bitwise
//     reproducible) and applies alpha / beta. With one split the epilogue is
//     fused.
//   * workgroup -> tile mapping is XCD aware: consecutive workgroups are
//     dispatched round-robin over the 8 XCDs, so the index is remapped such
//     that each XCD works on a compact group of tiles that share A and B
//     panels in its L2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels/gemm_desc.hpp"
#include "kernels/launch.hpp"
#include "kernels/tile_mma.hpp"

namespace spla_amd {
namespace {

using namespace tile;

__device__ __forceinline__ bool tile_active(const GemmDesc& d, int64_t r, int64_t c, int64_t nr,
                                            int64_t nc) {
  // SPLA block_is_active (src/block_generation/block.hpp:58-66) on a tile.
  if (d.fillMode == FillUpper) return c + nc + d.fillColOff > r + d.fillRowOff;
  if (d.fillMode == FillLower) return r + nr + d.fillRowOff > c + d.fillColOff;
  return true;
}

template <bool CPLX>
__device__ __forceinline__ void store_result(const GemmDesc& d, int i, int j, double re,
                                             double im) {
  if (CPLX) {
    double* c = d.C + (static_cast<int64_t>(j) * d.ldc + i) * 2;
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
    double* c = d.C + static_cast<int64_t>(j) * d.ldc + i;
    double out = d.alphaRe * re;
    if (!d.betaZero) out += d.betaRe * c[0];
    c[0] = out;
  }
}

// Maps the workgroup index to (split, tile row, tile column).
__device__ __forceinline__ void map_block(const GemmDesc& d, int& s, int& tm, int& tn) {
  const int G = gridDim.x;
  const int bid = blockIdx.x;
  int lin = bid;
  if ((G & 7) == 0) lin = (bid & 7) * (G >> 3) + (bid >> 3);
  const int tilesPerSplit = d.tilesM * d.tilesN;
  s = lin / tilesPerSplit;
  const int t = lin - s * tilesPerSplit;
  constexpr int kGroup = 8;
  const int group = t / (kGroup * d.tilesN);
  const int firstM = group * kGroup;
  const int gsize = min(d.tilesM - firstM, kGroup);
  const int inGroup = t - group * kGroup * d.tilesN;
  tm = firstM + inGroup % gsize;
  tn = inGroup / gsize;
}

template <bool CPLX, int LA, int LB, int WM, int WN, int WAVES_M, int BK, int MINBLOCKS>
__global__ void __launch_bounds__(kThreads, MINBLOCKS) gemm_f64_kernel(GemmDesc d) {
  using T = TileMma<CPLX, LA, LB, WM, WN, WAVES_M, BK>;
  __shared__ typename T::Smem sm;

  int s, tm, tn;
  map_block(d, s, tm, tn);
  const int row0 = tm * T::BM;  // embedded rows
  const int col0 = tn * T::BN;
  const int arow0 = CPLX ? row0 / 2 : row0;
  if (d.fillMode != FillFull) {
    const int64_t nr = min(T::AR, d.m - arow0);
    const int64_t nc = min(T::BN, d.n - col0);
    if (!tile_active(d, arow0, col0, nr, nc)) return;
  }
  const int64_t kBegin = static_cast<int64_t>(s) * d.kSplit;
  const int64_t kEnd = min(d.Kp, kBegin + d.kSplit);
  const TileOperands op{d.A, d.B, d.lda, d.ldb, d.m, d.n, d.conjA, d.conjB};
  dbl4 acc[T::FM][T::FN];
  T::run(op, row0, col0, kBegin, kEnd, sm, acc);
  if (d.splits > 1) {
    double* w = d.work + static_cast<int64_t>(s) * d.n * d.Mp;
    T::for_each(row0, col0, d.m, d.n, acc,
                [&](int i, int j, double re, double im) { store_work<CPLX>(w, d.Mp, i, j, re, im); });
  } else {
    T::for_each(row0, col0, d.m, d.n, acc,
                [&](int i, int j, double re, double im) { store_result<CPLX>(d, i, j, re, im); });
  }
}

// Adds the split-K partial sums in split order and applies alpha / beta.
template <bool CPLX>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(GemmDesc d, int tileRows,
                                                            int tileCols) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  if (idx >= total) return;
  const int i = static_cast<int>(idx % d.m);
  const int j = static_cast<int>(idx / d.m);
  if (d.fillMode != FillFull) {
    const int64_t r0 = (i / tileRows) * static_cast<int64_t>(tileRows);
    const int64_t c0 = (j / tileCols) * static_cast<int64_t>(tileCols);
    const int64_t nr = min<int64_t>(tileRows, d.m - r0);
    const int64_t nc = min<int64_t>(tileCols, d.n - c0);
    if (!tile_active(d, r0, c0, nr, nc)) return;
  }
  double re = 0.0, im = 0.0;
  const int64_t stride = static_cast<int64_t>(d.n) * d.Mp;
  if (CPLX) {
    const double* w = d.work + static_cast<int64_t>(j) * d.Mp + 2 * i;
    for (int s = 0; s < d.splits; ++s) {
      re += w[s * stride];
      im += w[s * stride + 1];
    }
  } else {
    const double* w = d.work + static_cast<int64_t>(j) * d.Mp + i;
    for (int s = 0; s < d.splits; ++s) re += w[s * stride];
  }
  store_result<CPLX>(d, i, j, re, im);
}

// C <- beta * C (k == 0 or alpha == 0); beta == 0 writes zeros without reading C.
template <bool CPLX>
__global__ void __launch_bounds__(256) scale_kernel(GemmDesc d) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  if (idx >= total) return;
  const int i = static_cast<int>(idx % d.m);
  const int j = static_cast<int>(idx / d.m);
  if (CPLX) {
    double* c = d.C + (static_cast<int64_t>(j) * d.ldc + i) * 2;
    if (d.betaZero) {
      c[0] = 0.0;
      c[1] = 0.0;
    } else {
      const double cr = c[0], ci = c[1];
      c[0] = d.betaRe * cr - d.betaIm * ci;
      c[1] = d.betaRe * ci + d.betaIm * cr;
    }
  } else {
    double* c = d.C + static_cast<int64_t>(j) * d.ldc + i;
    c[0] = d.betaZero ? 0.0 : d.betaRe * c[0];
  }
}

template <bool CPLX, int LA, int LB>
hipError_t launch_layout(const GemmDesc& d, int cfg, hipStream_t stream) {
  const dim3 block(kThreads);
  const dim3 grid(static_cast<unsigned>(d.tilesM * d.tilesN * d.splits));
  switch (cfg) {
    case TileM:
      hipLaunchKernelGGL((gemm_f64_kernel<CPLX, LA, LB, 32, 32, 2, 16, 2>), grid, block, 0, stream,
                         d);
      break;
    case TileXL:
      hipLaunchKernelGGL((gemm_f64_kernel<CPLX, LA, LB, 128, 64, 2, 16, 1>), grid, block, 0,
                         stream, d);
      break;
    default:
      hipLaunchKernelGGL((gemm_f64_kernel<CPLX, LA, LB, 64, 64, 2, 16, 2>), grid, block, 0, stream,
                         d);
      break;
  }
  return hipGetLastError();
}

template <bool CPLX>
hipError_t launch_typed(const GemmDesc& d, int la, int lb, int cfg, hipStream_t stream) {
  if (la == KMajor && lb == KMajor) return launch_layout<CPLX, KMajor, KMajor>(d, cfg, stream);
  if (la == KMajor && lb == MNMajor) return launch_layout<CPLX, KMajor, MNMajor>(d, cfg, stream);
  if (la == MNMajor && lb == KMajor) return launch_layout<CPLX, MNMajor, KMajor>(d, cfg, stream);
  return launch_layout<CPLX, MNMajor, MNMajor>(d, cfg, stream);
}

}  // namespace

hipError_t launch_gemm_f64(const GemmDesc& d, bool complex, int layoutA, int layoutB, int cfg,
                           hipStream_t stream) {
  return complex ? launch_typed<true>(d, layoutA, layoutB, cfg, stream)
                 : launch_typed<false>(d, layoutA, layoutB, cfg, stream);
}

hipError_t launch_splitk_reduce(const GemmDesc& d, bool complex, int cfg, hipStream_t stream) {
  const TileShape ts = tile_shape(cfg);
  const int tileRows = complex ? ts.bm / 2 : ts.bm;
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  const dim3 grid(static_cast<unsigned>((total + 255) / 256));
  if (complex)
    hipLaunchKernelGGL(splitk_reduce_kernel<true>, grid, dim3(256), 0, stream, d, tileRows, ts.bn);
  else
    hipLaunchKernelGGL(splitk_reduce_kernel<false>, grid, dim3(256), 0, stream, d, tileRows,
                       ts.bn);
  return hipGetLastError();
}

hipError_t launch_scale(const GemmDesc& d, bool complex, hipStream_t stream) {
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  if (total == 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((total + 255) / 256));
  if (complex)
    hipLaunchKernelGGL(scale_kernel<true>, grid, dim3(256), 0, stream, d);
  else
    hipLaunchKernelGGL(scale_kernel<false>, grid, dim3(256), 0, stream, d);
  return hipGetLastError();
}

}  // namespace spla_amd
