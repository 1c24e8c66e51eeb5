// Streaming FP64 / complex-FP64 A^T B and A^H B for small m, n and long
// reductions (both operands K-major: the ssb local product with opA = T or C,
// opB = N). Such products are bound by HBM bandwidth: a complex product moves
// 16 (m + n) bytes per reduction index for 8 m n flops, i.e. m n / (2 (m + n))
// flop per byte, below the FP64 matrix ridge (78.6 TFLOP/s / 8 TB/s = 9.8) for
// m = n <= 39.
//
// Each wave owns the whole (embedded) m x n output in FM x FN MFMA fragments
// and streams a share of the reduction straight from global memory into the
// MFMA operand registers: no LDS staging and no barrier per k-block. A wave
// step covers 16 embedded reduction indices, i.e. one 128-byte line of every
// row of A and B it reads; DEPTH steps of loads are in flight per wave
// (register ring), so that the CU keeps ~100-200 KB of loads outstanding.
// The four waves of a workgroup interleave their steps over the workgroup's
// reduction range and add their accumulators through LDS in wave order at the
// end; the workgroup's sum is one split-K partial, which the fixed-order
// reduction kernel adds to the others with alpha / beta.
//
// Fragment layout (v_mfma_f64_16x16x4_f64): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; the accumulator holds rows l / 16 + 4 r of column l % 16.
// Complex products use the real embedding of the tiled kernels: fragment rows
// 0-7 are the real and 8-15 the imaginary parts of 8 complex rows; a lane reads
// the (re, im) pair of one element and the imaginary rows swap it and flip
// signs lane-constantly (conjugation of A and / or B).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels/gemm_desc.hpp"
#include "kernels/launch.hpp"
#include "kernels/tile_mma.hpp"

namespace spla_amd {
namespace {

using namespace tile;

constexpr int kWaves = 4;
constexpr int kStep = 16;  // embedded reduction indices per wave step

// Two waves per SIMD (at most 256 registers per lane): with one, the m = n = 32
// complex product streamed 8 % less (3.57 vs 3.92 TB/s), its MFMAs and loads
// waiting on each other in the single wave.
template <bool CPLX, int FM, int FN, int DEPTH>
__global__ void __launch_bounds__(kWaves * 64, 2) skinny_gemm_kernel(GemmDesc d) {
  constexpr int ROWS = CPLX ? 8 : 16;  // rows of A per fragment (complex: re / im pairs)
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int l16 = lane & 15;
  const int kk = lane >> 4;
  const bool imagRow = CPLX && l16 >= 8;
  const int arow = CPLX ? (l16 & 7) : l16;
  uint32_t sgn0 = 0, sgn1 = 0;
  if (CPLX) {
    if (!imagRow) {
      sgn1 = (d.conjA == d.conjB) ? 0x80000000u : 0u;
    } else {
      sgn0 = d.conjA ? 0x80000000u : 0u;
      sgn1 = d.conjB ? 0x80000000u : 0u;
    }
  }
  const double* __restrict__ A = static_cast<const double*>(d.A);
  const double* __restrict__ B = static_cast<const double*>(d.B);
  const int64_t comps = CPLX ? 2 : 1;
  const int64_t kBegin = static_cast<int64_t>(blockIdx.x) * d.kSplit;
  const int64_t kEnd = min(d.Kp, kBegin + d.kSplit);
  const int64_t steps = (kEnd - kBegin + kStep - 1) / kStep;
  // this wave's steps: wave, wave + 4, wave + 8, ...
  const int64_t mySteps = steps > wave ? (steps - wave + kWaves - 1) / kWaves : 0;

  // row / column bases of this lane; rows past m and columns past n read row /
  // column 0 and are zeroed after the load, so that every load is issued
  // unconditionally and the waits for the register ring stay static
  const double* pa[FM];
  bool va[FM];
#pragma unroll
  for (int f = 0; f < FM; ++f) {
    const int r = ROWS * f + arow;
    va[f] = r < d.m;
    pa[f] = A + static_cast<int64_t>(va[f] ? r : 0) * d.lda * comps;
  }
  const double* pb[FN];
  bool vb[FN];
#pragma unroll
  for (int g = 0; g < FN; ++g) {
    const int c = 16 * g + l16;
    vb[g] = c < d.n;
    pb[g] = B + static_cast<int64_t>(vb[g] ? c : 0) * d.ldb * comps;
  }

  typedef double dbl2v __attribute__((ext_vector_type(2)));
  dbl2v ra[DEPTH][FM][2], rb[DEPTH][FN][2];
  dbl4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; ++f)
#pragma unroll
    for (int g = 0; g < FN; ++g) acc[f][g] = dbl4{0.0, 0.0, 0.0, 0.0};

  // loads of wave step t into a ring slot; sub-block h of the step holds
  // embedded indices q0 + 8 h + 2 kk and q0 + 8 h + 2 kk + 1. Indices past kEnd
  // (the range's tail, and the steps past the wave's last one that the ring's
  // tail loads) read the range's first element, rows past m and columns past n
  // read row / column 0; compute() zeroes all of them where it uses the values
  // (zeroing right after the load would make the wave wait for each load), so
  // the ring runs without a branch and the surplus steps add zeros.
  auto step_q = [&](int64_t t, int h) {
    return kBegin + (wave + t * kWaves) * static_cast<int64_t>(kStep) + 8 * h + 2 * kk;
  };
  auto load = [&](int64_t t, auto& sa, auto& sb) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t q = step_q(t, h);
      const bool in0 = q < kEnd, in1 = q + 1 < kEnd;  // complex: in1 == in0 (q, kEnd even)
      const int64_t qc = in0 ? q : kBegin;            // a valid address for every lane
      const int64_t qc1 = in1 ? q + 1 : qc;
#pragma unroll
      for (int f = 0; f < FM; ++f) {
        if (CPLX) {
          sa[f][h] = *reinterpret_cast<const dbl2v*>(pa[f] + qc);
        } else {
          sa[f][h].x = pa[f][qc];
          sa[f][h].y = pa[f][qc1];
        }
      }
#pragma unroll
      for (int g = 0; g < FN; ++g) {
        if (CPLX) {
          sb[g][h] = *reinterpret_cast<const dbl2v*>(pb[g] + qc);
        } else {
          sb[g][h].x = pb[g][qc];
          sb[g][h].y = pb[g][qc1];
        }
      }
    }
  };
  auto compute = [&](int64_t t, const auto& sa, const auto& sb) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t q = step_q(t, h);
      const bool in0 = q < kEnd, in1 = q + 1 < kEnd;
      double a0[FM], a1[FM], b0[FN], b1[FN];
#pragma unroll
      for (int f = 0; f < FM; ++f) {
        const dbl2v v = sa[f][h];
        const double vx = (va[f] && in0) ? v.x : 0.0;
        const double vy = (va[f] && in1) ? v.y : 0.0;
        if (CPLX) {
          const double x = imagRow ? vy : vx;
          const double y = imagRow ? vx : vy;
          a0[f] = flip_sign(x, sgn0);
          a1[f] = flip_sign(y, sgn1);
        } else {
          a0[f] = vx;
          a1[f] = vy;
        }
      }
#pragma unroll
      for (int g = 0; g < FN; ++g) {
        b0[g] = (vb[g] && in0) ? sb[g][h].x : 0.0;
        b1[g] = (vb[g] && in1) ? sb[g][h].y : 0.0;
      }
#pragma unroll
      for (int f = 0; f < FM; ++f)
#pragma unroll
        for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a0[f], b0[g], acc[f][g]);
#pragma unroll
      for (int f = 0; f < FM; ++f)
#pragma unroll
        for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a1[f], b1[g], acc[f][g]);
    }
  };

  if (mySteps > 0) {
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) load(u, ra[u], rb[u]);
    for (int64_t base = 0; base < mySteps; base += DEPTH) {
#pragma unroll
      for (int u = 0; u < DEPTH; ++u) {
        compute(base + u, ra[u], rb[u]);
        load(base + u + DEPTH, ra[u], rb[u]);
      }
    }
  }

  // the four waves' sums, added in wave order
  __shared__ dbl4 part[kWaves - 1][FM][FN][64];
  if (wave > 0) {
#pragma unroll
    for (int f = 0; f < FM; ++f)
#pragma unroll
      for (int g = 0; g < FN; ++g) part[wave - 1][f][g][lane] = acc[f][g];
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int w = 0; w < kWaves - 1; ++w)
#pragma unroll
    for (int f = 0; f < FM; ++f)
#pragma unroll
      for (int g = 0; g < FN; ++g) acc[f][g] += part[w][f][g][lane];

  // accumulator rows kk + 4 r of column l16 (see tile_mma.hpp for_each)
  double* work = static_cast<double*>(d.work) + static_cast<int64_t>(blockIdx.x) * d.n * d.Mp;
#pragma unroll
  for (int f = 0; f < FM; ++f)
#pragma unroll
    for (int g = 0; g < FN; ++g) {
      const int j = 16 * g + l16;
      if (j >= d.n) continue;
      const dbl4 v = acc[f][g];
      if (CPLX) {
        const int base = 8 * f;
        if (base + kk < d.m) store_work<double, true>(work, d.Mp, base + kk, j, v[0], v[2]);
        if (base + kk + 4 < d.m) store_work<double, true>(work, d.Mp, base + kk + 4, j, v[1], v[3]);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (16 * f + kk + 4 * r < d.m)
            store_work<double, false>(work, d.Mp, 16 * f + kk + 4 * r, j, v[r], 0.0);
      }
    }
}

template <bool CPLX, int FM, int FN>
hipError_t launch_fm_fn(const GemmDesc& d, hipStream_t stream) {
  // loads in flight per wave: DEPTH steps of (FM + FN) x 2 x 16 bytes per lane
  // (the largest ring that fits two waves per SIMD; measured 2 over 3 for FM + FN > 4)
  constexpr int DEPTH = (FM + FN) <= 3 ? 6 : ((FM + FN) <= 4 ? 4 : 2);
  hipLaunchKernelGGL((skinny_gemm_kernel<CPLX, FM, FN, DEPTH>), dim3(d.splits),
                     dim3(kWaves * 64), 0, stream, d);
  return hipGetLastError();
}

template <bool CPLX>
hipError_t launch_typed(const GemmDesc& d, hipStream_t stream) {
  constexpr int ROWS = CPLX ? 8 : 16;
  const int fm = d.m <= ROWS ? 1 : (d.m <= 2 * ROWS ? 2 : 4);
  const int fn = d.n <= 16 ? 1 : 2;
  if (fm == 1) return fn == 1 ? launch_fm_fn<CPLX, 1, 1>(d, stream) : launch_fm_fn<CPLX, 1, 2>(d, stream);
  if (fm == 2) return fn == 1 ? launch_fm_fn<CPLX, 2, 1>(d, stream) : launch_fm_fn<CPLX, 2, 2>(d, stream);
  return fn == 1 ? launch_fm_fn<CPLX, 4, 1>(d, stream) : launch_fm_fn<CPLX, 4, 2>(d, stream);
}

}  // namespace

bool skinny_supported(bool complex, int layoutA, int layoutB, int64_t m, int64_t n) {
  const int64_t rows = complex ? 8 : 16;
  return layoutA == KMajor && layoutB == KMajor && m >= 1 && n >= 1 && m <= 4 * rows && n <= 32;
}

hipError_t launch_skinny_gemm(const GemmDesc& d, bool complex, hipStream_t stream) {
  return complex ? launch_typed<true>(d, stream) : launch_typed<false>(d, stream);
}

}  // namespace spla_amd
