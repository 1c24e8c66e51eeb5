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

namespace spla_amd {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

constexpr int kThreads = 256;
constexpr int kPad = 2;  // LDS row padding (doubles): keeps 16-byte alignment, spreads banks

__device__ __forceinline__ double flip_sign(double x, uint32_t hiMask) {
  uint64_t b = __builtin_bit_cast(uint64_t, x);
  b ^= static_cast<uint64_t>(hiMask) << 32;
  return __builtin_bit_cast(double, b);
}

// Stages one operand tile (ROWS stored rows x BK embedded reduction elements)
// from global memory into registers and from registers into LDS.
template <bool CPLX, int LAYOUT, int ROWS, int BK>
struct TileStage {
  static constexpr int kPairs = ROWS * BK / 2;
  static constexpr int kPerThread = kPairs / kThreads;
  static_assert(kPairs % kThreads == 0, "tile does not divide over the workgroup");
  dbl2 regs[kPerThread];

  __device__ __forceinline__ static void coords(int pi, int& row, int& qp) {
    if (LAYOUT == KMajor) {
      row = pi / (BK / 2);
      qp = pi % (BK / 2);
    } else {
      row = pi % ROWS;
      qp = pi / ROWS;
    }
  }

  // base: operand origin (doubles); ld: leading dimension in elements;
  // rowBase/rowLimit: stored-row range; q0/qEnd: embedded reduction range.
  __device__ __forceinline__ void load(const double* __restrict__ base, int64_t ld, int rowBase,
                                       int rowLimit, int64_t q0, int64_t qEnd, int tid) {
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      int row, qp;
      coords(tid + u * kThreads, row, qp);
      const int grow = rowBase + row;
      const int64_t q = q0 + 2 * qp;
      double d0 = 0.0, d1 = 0.0;
      if (grow < rowLimit) {
        if (LAYOUT == KMajor) {
          const double* p = base + static_cast<int64_t>(grow) * ld * (CPLX ? 2 : 1) + q;
          if (q < qEnd) d0 = p[0];
          if (q + 1 < qEnd) d1 = p[1];
        } else if (CPLX) {
          if (q < qEnd) {
            const double* p = base + ((q >> 1) * ld + grow) * 2;
            d0 = p[0];
            d1 = p[1];
          }
        } else {
          if (q < qEnd) d0 = base[q * ld + grow];
          if (q + 1 < qEnd) d1 = base[(q + 1) * ld + grow];
        }
      }
      regs[u] = dbl2{d0, d1};
    }
  }

  __device__ __forceinline__ void store(double (*lds)[BK + kPad], int tid) const {
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      int row, qp;
      coords(tid + u * kThreads, row, qp);
      *reinterpret_cast<dbl2*>(&lds[row][2 * qp]) = regs[u];
    }
  }
};

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
  constexpr int WAVES_N = 4 / WAVES_M;
  constexpr int BM = WM * WAVES_M;
  constexpr int BN = WN * WAVES_N;
  constexpr int AR = CPLX ? BM / 2 : BM;
  constexpr int FM = WM / 16;
  constexpr int FN = WN / 16;
  static_assert(BK % 8 == 0, "k-block must hold whole 8-element sub-blocks");

  __shared__ double lA[2][AR][BK + kPad];
  __shared__ double lB[2][BN][BK + kPad];

  int s, tm, tn;
  map_block(d, s, tm, tn);
  const int row0 = tm * BM;  // embedded rows
  const int col0 = tn * BN;
  const int arow0 = CPLX ? row0 / 2 : row0;

  if (d.fillMode != FillFull) {
    const int64_t nr = min(AR, d.m - arow0);
    const int64_t nc = min(BN, d.n - col0);
    if (!tile_active(d, arow0, col0, nr, nc)) return;
  }

  const int64_t kBegin = static_cast<int64_t>(s) * d.kSplit;
  const int64_t kEnd = min(d.Kp, kBegin + d.kSplit);
  const int nk = static_cast<int>((kEnd - kBegin + BK - 1) / BK);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wm = wid % WAVES_M;
  const int wn = wid / WAVES_M;
  const int l16 = lane & 15;
  const int kk = lane >> 4;

  // Lane-constant complex operand selection (see gemm_desc.hpp):
  //   real row  (t = 0): step0 <- Re(a),  step1 <- Im(a) * (conjA == conjB ? -1 : 1)
  //   imag row  (t = 1): step0 <- Im(a) * (conjA ? -1 : 1), step1 <- Re(a) * (conjB ? -1 : 1)
  // Fragment row r = lane % 16 is an imaginary row for r >= 8.
  bool swapPair = false;
  uint32_t sgn0 = 0, sgn1 = 0;
  if (CPLX) {
    swapPair = l16 >= 8;
    if (!swapPair) {
      sgn1 = (d.conjA == d.conjB) ? 0x80000000u : 0u;
    } else {
      sgn0 = d.conjA ? 0x80000000u : 0u;
      sgn1 = d.conjB ? 0x80000000u : 0u;
    }
  }

  dbl4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; ++f)
#pragma unroll
    for (int g = 0; g < FN; ++g) acc[f][g] = dbl4{0.0, 0.0, 0.0, 0.0};

  TileStage<CPLX, LA, AR, BK> stA;
  TileStage<CPLX, LB, BN, BK> stB;

  if (nk > 0) {
    stA.load(d.A, d.lda, arow0, d.m, kBegin, kEnd, tid);
    stB.load(d.B, d.ldb, col0, d.n, kBegin, kEnd, tid);
    stA.store(lA[0], tid);
    stB.store(lB[0], tid);
  }
  __syncthreads();

  for (int it = 0; it < nk; ++it) {
    const int cur = it & 1;
    const bool more = it + 1 < nk;
    if (more) {
      const int64_t q0 = kBegin + static_cast<int64_t>(it + 1) * BK;
      stA.load(d.A, d.lda, arow0, d.m, q0, kEnd, tid);
      stB.load(d.B, d.ldb, col0, d.n, q0, kEnd, tid);
    }

#pragma unroll
    for (int kb8 = 0; kb8 < BK / 8; ++kb8) {
      double a0[FM], a1[FM];
      dbl2 b[FN];
#pragma unroll
      for (int f = 0; f < FM; ++f) {
        const int sr = CPLX ? (wm * WM) / 2 + 8 * f + (l16 & 7) : wm * WM + 16 * f + l16;
        const dbl2 v = *reinterpret_cast<const dbl2*>(&lA[cur][sr][kb8 * 8 + 2 * kk]);
        if (CPLX) {
          const double x = swapPair ? v.y : v.x;
          const double y = swapPair ? v.x : v.y;
          a0[f] = flip_sign(x, sgn0);
          a1[f] = flip_sign(y, sgn1);
        } else {
          a0[f] = v.x;
          a1[f] = v.y;
        }
      }
#pragma unroll
      for (int g = 0; g < FN; ++g) {
        const int j = wn * WN + 16 * g + l16;
        b[g] = *reinterpret_cast<const dbl2*>(&lB[cur][j][kb8 * 8 + 2 * kk]);
      }
#pragma unroll
      for (int f = 0; f < FM; ++f)
#pragma unroll
        for (int g = 0; g < FN; ++g)
          acc[f][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[f], b[g].x, acc[f][g], 0, 0, 0);
#pragma unroll
      for (int f = 0; f < FM; ++f)
#pragma unroll
        for (int g = 0; g < FN; ++g)
          acc[f][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[f], b[g].y, acc[f][g], 0, 0, 0);
    }

    if (more) {
      stA.store(lA[cur ^ 1], tid);
      stB.store(lB[cur ^ 1], tid);
    }
    __syncthreads();
  }

  // Accumulator layout of v_mfma_f64_16x16x4_f64: lane holds rows kk + 4 * r,
  // r = 0..3, of column lane % 16 (kk = lane / 16). For complex fragments this
  // is (re, im) of complex rows kk and kk + 4 in (acc[0], acc[2]) and
  // (acc[1], acc[3]).
#pragma unroll
  for (int f = 0; f < FM; ++f) {
#pragma unroll
    for (int g = 0; g < FN; ++g) {
      const int j = col0 + wn * WN + 16 * g + l16;
      if (j >= d.n) continue;
      const dbl4 v = acc[f][g];
      if (CPLX) {
        const int i0 = arow0 + (wm * WM) / 2 + 8 * f + kk;
        if (d.splits > 1) {
          double* w = d.work + (static_cast<int64_t>(s) * d.n + j) * d.Mp;
          if (i0 < d.m) {
            w[2 * i0] = v[0];
            w[2 * i0 + 1] = v[2];
          }
          if (i0 + 4 < d.m) {
            w[2 * i0 + 8] = v[1];
            w[2 * i0 + 9] = v[3];
          }
        } else {
          if (i0 < d.m) store_result<true>(d, i0, j, v[0], v[2]);
          if (i0 + 4 < d.m) store_result<true>(d, i0 + 4, j, v[1], v[3]);
        }
      } else {
        const int p0 = row0 + wm * WM + 16 * f + kk;
        if (d.splits > 1) {
          double* w = d.work + (static_cast<int64_t>(s) * d.n + j) * d.Mp;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (p0 + 4 * r < d.m) w[p0 + 4 * r] = v[r];
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (p0 + 4 * r < d.m) store_result<false>(d, p0 + 4 * r, j, v[r], 0.0);
        }
      }
    }
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
