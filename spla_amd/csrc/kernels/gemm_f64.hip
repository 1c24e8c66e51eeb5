// FP64 / complex-FP64 GEMM on the gfx950 FP64 matrix cores (v_mfma_f64_16x16x4_f64),
// and the same kernels in FP32 / complex-FP32 on the FP32 matrix cores
// (v_mfma_f32_16x16x4_f32, FP32 accumulation as in a single-precision BLAS).
//
// Replaces the vendor ?gemm call of SPLA's local engine
// (src/gpu_util/multiply_gpu.cpp:40-76 -> rocblas_{d,z}gemm) for the
// operations the engine issues: opA in {N, T, C}, opB in {N, T, C}.
//
// Structure of one workgroup (256 threads = 4 wave64):
//   * a BM x BN tile of the embedded output (see gemm_desc.hpp), split over a
//     WAVES_M x WAVES_N grid of waves, each owning WM x WN = (WM/16)x(WN/16)
//     MFMA accumulators of 16x16 scalars;
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

// C <- alpha * (re, im) + beta * C in the scalar type S (beta == 0: C not read).
template <typename S, bool CPLX>
__device__ __forceinline__ void store_result(const GemmDesc& d, int i, int j, S re, S im) {
  const S aRe = static_cast<S>(d.alphaRe), aIm = static_cast<S>(d.alphaIm);
  const S bRe = static_cast<S>(d.betaRe), bIm = static_cast<S>(d.betaIm);
  if (CPLX) {
    S* c = static_cast<S*>(d.C) + (static_cast<int64_t>(j) * d.ldc + i) * 2;
    S outRe = aRe * re - aIm * im;
    S outIm = aRe * im + aIm * re;
    if (!d.betaZero) {
      const S cr = c[0], ci = c[1];
      outRe += bRe * cr - bIm * ci;
      outIm += bRe * ci + bIm * cr;
    }
    c[0] = outRe;
    c[1] = outIm;
  } else {
    S* c = static_cast<S*>(d.C) + static_cast<int64_t>(j) * d.ldc + i;
    S out = aRe * re;
    if (!d.betaZero) out += bRe * c[0];
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

// Body of one workgroup (shared by the FP64 and FP32 kernels).
template <typename S, bool CPLX, int LA, int LB, int WM, int WN, int WAVES_M, int BK>
__device__ __forceinline__ void gemm_tile(const GemmDesc& d,
                                          typename TileMma<S, CPLX, LA, LB, WM, WN, WAVES_M,
                                                           BK>::Smem& sm) {
  using T = TileMma<S, CPLX, LA, LB, WM, WN, WAVES_M, BK>;
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
  const TileOperands<S> op{static_cast<const S*>(d.A), static_cast<const S*>(d.B), d.lda, d.ldb,
                           d.m, d.n, d.conjA, d.conjB};
  typename T::Acc acc[T::FM][T::FN];
  T::run(op, row0, col0, kBegin, kEnd, sm, acc);
  if (d.splits > 1) {
    S* w = static_cast<S*>(d.work) + static_cast<int64_t>(s) * d.n * d.Mp;
    T::for_each(row0, col0, d.m, d.n, acc,
                [&](int i, int j, S re, S im) { store_work<S, CPLX>(w, d.Mp, i, j, re, im); });
  } else {
    T::for_each(row0, col0, d.m, d.n, acc,
                [&](int i, int j, S re, S im) { store_result<S, CPLX>(d, i, j, re, im); });
  }
}

template <bool CPLX, int LA, int LB, int WM, int WN, int WAVES_M, int BK, int MINBLOCKS>
__global__ void __launch_bounds__(kThreads, MINBLOCKS) gemm_f64_kernel(GemmDesc d) {
  __shared__ typename TileMma<double, CPLX, LA, LB, WM, WN, WAVES_M, BK>::Smem sm;
  gemm_tile<double, CPLX, LA, LB, WM, WN, WAVES_M, BK>(d, sm);
}

template <bool CPLX, int LA, int LB, int WM, int WN, int WAVES_M, int BK, int MINBLOCKS>
__global__ void __launch_bounds__(kThreads, MINBLOCKS) gemm_f32_kernel(GemmDesc d) {
  __shared__ typename TileMma<float, CPLX, LA, LB, WM, WN, WAVES_M, BK>::Smem sm;
  gemm_tile<float, CPLX, LA, LB, WM, WN, WAVES_M, BK>(d, sm);
}

// Adds the split-K partial sums in split order and applies alpha / beta.
template <typename S, bool CPLX>
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
  S re = S(0), im = S(0);
  const int64_t stride = static_cast<int64_t>(d.n) * d.Mp;
  if (CPLX) {
    const S* w = static_cast<const S*>(d.work) + static_cast<int64_t>(j) * d.Mp + 2 * i;
    for (int s = 0; s < d.splits; ++s) {
      re += w[s * stride];
      im += w[s * stride + 1];
    }
  } else {
    const S* w = static_cast<const S*>(d.work) + static_cast<int64_t>(j) * d.Mp + i;
    for (int s = 0; s < d.splits; ++s) re += w[s * stride];
  }
  store_result<S, CPLX>(d, i, j, re, im);
}

// Many splits over few elements (small m, n, long reduction): 16 elements per
// block, 16 split groups per element. Group g sums splits g, g + 16, g + 32, ...
// in order and the 16 group sums are added in group order, so the result is
// bitwise reproducible; every split is read by 16 consecutive elements at once.
constexpr int kWideElems = 16;
constexpr int kWideGroups = 16;
template <typename S, bool CPLX>
__global__ void __launch_bounds__(256) splitk_reduce_wide_kernel(GemmDesc d, int tileRows,
                                                                 int tileCols) {
  __shared__ S part[kWideGroups][kWideElems][2];
  const int e = threadIdx.x % kWideElems;
  const int g = threadIdx.x / kWideElems;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kWideElems + e;
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  bool live = idx < total;
  const int i = live ? static_cast<int>(idx % d.m) : 0;
  const int j = live ? static_cast<int>(idx / d.m) : 0;
  if (live && d.fillMode != FillFull) {
    const int64_t r0 = (i / tileRows) * static_cast<int64_t>(tileRows);
    const int64_t c0 = (j / tileCols) * static_cast<int64_t>(tileCols);
    const int64_t nr = min<int64_t>(tileRows, d.m - r0);
    const int64_t nc = min<int64_t>(tileCols, d.n - c0);
    live = tile_active(d, r0, c0, nr, nc);
  }
  S re = S(0), im = S(0);
  if (live) {
    const int64_t stride = static_cast<int64_t>(d.n) * d.Mp;
    const S* w = static_cast<const S*>(d.work) + static_cast<int64_t>(j) * d.Mp + (CPLX ? 2 * i : i);
    for (int sp = g; sp < d.splits; sp += kWideGroups) {
      re += w[sp * stride];
      if (CPLX) im += w[sp * stride + 1];
    }
  }
  part[g][e][0] = re;
  part[g][e][1] = im;
  __syncthreads();
  if (g == 0 && live) {
    S sr = part[0][e][0], si = part[0][e][1];
#pragma unroll
    for (int q = 1; q < kWideGroups; ++q) {
      sr += part[q][e][0];
      si += part[q][e][1];
    }
    store_result<S, CPLX>(d, i, j, sr, si);
  }
}

// C <- beta * C (k == 0 or alpha == 0); beta == 0 writes zeros without reading C.
template <typename S, bool CPLX>
__global__ void __launch_bounds__(256) scale_kernel(GemmDesc d) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  if (idx >= total) return;
  const int i = static_cast<int>(idx % d.m);
  const int j = static_cast<int>(idx / d.m);
  const S bRe = static_cast<S>(d.betaRe), bIm = static_cast<S>(d.betaIm);
  if (CPLX) {
    S* c = static_cast<S*>(d.C) + (static_cast<int64_t>(j) * d.ldc + i) * 2;
    if (d.betaZero) {
      c[0] = S(0);
      c[1] = S(0);
    } else {
      const S cr = c[0], ci = c[1];
      c[0] = bRe * cr - bIm * ci;
      c[1] = bRe * ci + bIm * cr;
    }
  } else {
    S* c = static_cast<S*>(d.C) + static_cast<int64_t>(j) * d.ldc + i;
    c[0] = d.betaZero ? S(0) : bRe * c[0];
  }
}

template <bool F32, bool CPLX, int LA, int LB>
hipError_t launch_layout(const GemmDesc& d, int cfg, hipStream_t stream) {
  const dim3 block(kThreads);
  const dim3 grid(static_cast<unsigned>(d.tilesM * d.tilesN * d.splits));
  if (F32) {
    switch (cfg) {
      case TileM:
        hipLaunchKernelGGL((gemm_f32_kernel<CPLX, LA, LB, 32, 32, 2, 16, 2>), grid, block, 0,
                           stream, d);
        break;
      case TileXL:
        hipLaunchKernelGGL((gemm_f32_kernel<CPLX, LA, LB, 128, 64, 2, 16, 1>), grid, block, 0,
                           stream, d);
        break;
      default:
        hipLaunchKernelGGL((gemm_f32_kernel<CPLX, LA, LB, 64, 64, 2, 16, 2>), grid, block, 0,
                           stream, d);
        break;
    }
    return hipGetLastError();
  }
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

template <bool F32, bool CPLX>
hipError_t launch_typed(const GemmDesc& d, int la, int lb, int cfg, hipStream_t stream) {
  if (la == KMajor && lb == KMajor) return launch_layout<F32, CPLX, KMajor, KMajor>(d, cfg, stream);
  if (la == KMajor && lb == MNMajor)
    return launch_layout<F32, CPLX, KMajor, MNMajor>(d, cfg, stream);
  if (la == MNMajor && lb == KMajor)
    return launch_layout<F32, CPLX, MNMajor, KMajor>(d, cfg, stream);
  return launch_layout<F32, CPLX, MNMajor, MNMajor>(d, cfg, stream);
}

}  // namespace

hipError_t launch_gemm_f64(const GemmDesc& d, bool complex, int layoutA, int layoutB, int cfg,
                           hipStream_t stream) {
  return complex ? launch_typed<false, true>(d, layoutA, layoutB, cfg, stream)
                 : launch_typed<false, false>(d, layoutA, layoutB, cfg, stream);
}

hipError_t launch_gemm_f32(const GemmDesc& d, bool complex, int layoutA, int layoutB, int cfg,
                           hipStream_t stream) {
  return complex ? launch_typed<true, true>(d, layoutA, layoutB, cfg, stream)
                 : launch_typed<true, false>(d, layoutA, layoutB, cfg, stream);
}

template <typename S>
hipError_t launch_splitk_reduce_s(const GemmDesc& d, bool complex, int cfg, hipStream_t stream) {
  const TileShape ts = tile_shape(cfg);
  const int tileRows = complex ? ts.bm / 2 : ts.bm;
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  if (d.splits > 64) {  // only the whole-device plans of small m, n split this far
    const dim3 wide(static_cast<unsigned>((total + kWideElems - 1) / kWideElems));
    if (complex)
      hipLaunchKernelGGL((splitk_reduce_wide_kernel<S, true>), wide, dim3(256), 0, stream, d,
                         tileRows, ts.bn);
    else
      hipLaunchKernelGGL((splitk_reduce_wide_kernel<S, false>), wide, dim3(256), 0, stream, d,
                         tileRows, ts.bn);
    return hipGetLastError();
  }
  const dim3 grid(static_cast<unsigned>((total + 255) / 256));
  if (complex)
    hipLaunchKernelGGL((splitk_reduce_kernel<S, true>), grid, dim3(256), 0, stream, d, tileRows,
                       ts.bn);
  else
    hipLaunchKernelGGL((splitk_reduce_kernel<S, false>), grid, dim3(256), 0, stream, d, tileRows,
                       ts.bn);
  return hipGetLastError();
}

hipError_t launch_splitk_reduce(const GemmDesc& d, bool complex, int cfg, hipStream_t stream,
                                bool f32) {
  return f32 ? launch_splitk_reduce_s<float>(d, complex, cfg, stream)
             : launch_splitk_reduce_s<double>(d, complex, cfg, stream);
}

hipError_t launch_scale(const GemmDesc& d, bool complex, hipStream_t stream, bool f32) {
  const int64_t total = static_cast<int64_t>(d.m) * d.n;
  if (total == 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((total + 255) / 256));
  if (f32) {
    if (complex)
      hipLaunchKernelGGL((scale_kernel<float, true>), grid, dim3(256), 0, stream, d);
    else
      hipLaunchKernelGGL((scale_kernel<float, false>), grid, dim3(256), 0, stream, d);
  } else {
    if (complex)
      hipLaunchKernelGGL((scale_kernel<double, true>), grid, dim3(256), 0, stream, d);
    else
      hipLaunchKernelGGL((scale_kernel<double, false>), grid, dim3(256), 0, stream, d);
  }
  return hipGetLastError();
}

}  // namespace spla_amd
