// FP64 MFMA tile routine shared by the single-product kernel (gemm_f64.hip)
// and the persistent batched kernel (batched.hip). Device code only; include
// from HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels/gemm_desc.hpp"

namespace spla_amd {
namespace tile {

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
// UNCHECKED: k-blocks wholly inside the operands skip the per-element checks
// (measured faster on the large tiles, slower on the small ones).
template <bool CPLX, int LAYOUT, int ROWS, int BK, bool UNCHECKED>
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
  // A k-block wholly inside both ranges is loaded without per-element checks.
  __device__ __forceinline__ void load(const double* __restrict__ base, int64_t ld, int rowBase,
                                       int rowLimit, int64_t q0, int64_t qEnd, int tid) {
    if (UNCHECKED && rowBase + ROWS <= rowLimit && q0 + BK <= qEnd) {
#pragma unroll
      for (int u = 0; u < kPerThread; ++u) {
        int row, qp;
        coords(tid + u * kThreads, row, qp);
        const int grow = rowBase + row;
        const int64_t q = q0 + 2 * qp;
        if (LAYOUT == KMajor) {
          const double* p = base + static_cast<int64_t>(grow) * ld * (CPLX ? 2 : 1) + q;
          regs[u] = dbl2{p[0], p[1]};
        } else if (CPLX) {
          const double* p = base + ((q >> 1) * ld + grow) * 2;
          regs[u] = dbl2{p[0], p[1]};
        } else {
          regs[u] = dbl2{base[q * ld + grow], base[(q + 1) * ld + grow]};
        }
      }
      return;
    }
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

// Operands of one tile product (element rows m / columns n of C).
struct TileOperands {
  const double* A;
  const double* B;
  int64_t lda, ldb;
  int m, n;
  int conjA, conjB;
};

// One BM x BN tile of the embedded product over the embedded reduction range
// [kBegin, kEnd), accumulated in registers. Shared by the single-product
// kernel and the persistent batched kernel. On return every thread has passed
// the final barrier, so the LDS buffers may be reused at once.
template <bool CPLX, int LA, int LB, int WM, int WN, int WAVES_M, int BK>
struct TileMma {
  static constexpr int WAVES_N = 4 / WAVES_M;
  static constexpr int BM = WM * WAVES_M;
  static constexpr int BN = WN * WAVES_N;
  static constexpr int AR = CPLX ? BM / 2 : BM;
  static constexpr int FM = WM / 16;
  static constexpr int FN = WN / 16;
  static_assert(BK % 8 == 0, "k-block must hold whole 8-element sub-blocks");

  struct Smem {
    double lA[2][AR][BK + kPad];
    double lB[2][BN][BK + kPad];
  };

  __device__ __forceinline__ static void run(const TileOperands& op, int row0, int col0,
                                             int64_t kBegin, int64_t kEnd, Smem& sm,
                                             dbl4 (&acc)[FM][FN]) {
    const int arow0 = CPLX ? row0 / 2 : row0;
    const int nk = static_cast<int>((kEnd - kBegin + BK - 1) / BK);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    const int wm = wid % WAVES_M;
    const int wn = wid / WAVES_M;
    const int l16 = lane & 15;
    const int kk = lane >> 4;

    // Lane-constant complex operand selection (see gemm_desc.hpp):
    //   real row (t = 0): step0 <- Re(a), step1 <- Im(a) * (conjA == conjB ? -1 : 1)
    //   imag row (t = 1): step0 <- Im(a) * (conjA ? -1 : 1), step1 <- Re(a) * (conjB ? -1 : 1)
    // Fragment row r = lane % 16 is an imaginary row for r >= 8.
    bool swapPair = false;
    uint32_t sgn0 = 0, sgn1 = 0;
    if (CPLX) {
      swapPair = l16 >= 8;
      if (!swapPair) {
        sgn1 = (op.conjA == op.conjB) ? 0x80000000u : 0u;
      } else {
        sgn0 = op.conjA ? 0x80000000u : 0u;
        sgn1 = op.conjB ? 0x80000000u : 0u;
      }
    }

#pragma unroll
    for (int f = 0; f < FM; ++f)
#pragma unroll
      for (int g = 0; g < FN; ++g) acc[f][g] = dbl4{0.0, 0.0, 0.0, 0.0};

    constexpr bool kUnchecked = WM * WN >= 64 * 64;
    TileStage<CPLX, LA, AR, BK, kUnchecked> stA;
    TileStage<CPLX, LB, BN, BK, kUnchecked> stB;

    if (nk > 0) {
      stA.load(op.A, op.lda, arow0, op.m, kBegin, kEnd, tid);
      stB.load(op.B, op.ldb, col0, op.n, kBegin, kEnd, tid);
      stA.store(sm.lA[0], tid);
      stB.store(sm.lB[0], tid);
    }
    __syncthreads();

    for (int it = 0; it < nk; ++it) {
      const int cur = it & 1;
      const bool more = it + 1 < nk;
      if (more) {
        const int64_t q0 = kBegin + static_cast<int64_t>(it + 1) * BK;
        stA.load(op.A, op.lda, arow0, op.m, q0, kEnd, tid);
        stB.load(op.B, op.ldb, col0, op.n, q0, kEnd, tid);
      }

#pragma unroll
      for (int kb8 = 0; kb8 < BK / 8; ++kb8) {
        double a0[FM], a1[FM];
        dbl2 b[FN];
#pragma unroll
        for (int f = 0; f < FM; ++f) {
          const int sr = CPLX ? (wm * WM) / 2 + 8 * f + (l16 & 7) : wm * WM + 16 * f + l16;
          const dbl2 v = *reinterpret_cast<const dbl2*>(&sm.lA[cur][sr][kb8 * 8 + 2 * kk]);
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
          b[g] = *reinterpret_cast<const dbl2*>(&sm.lB[cur][j][kb8 * 8 + 2 * kk]);
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
        stA.store(sm.lA[cur ^ 1], tid);
        stB.store(sm.lB[cur ^ 1], tid);
      }
      __syncthreads();
    }
  }

  // Calls f(i, j, re, im) for every element of the tile inside m x n.
  // Accumulator layout of v_mfma_f64_16x16x4_f64: lane holds rows kk + 4 * r,
  // r = 0..3, of column lane % 16 (kk = lane / 16). For complex fragments this
  // is (re, im) of complex rows kk and kk + 4 in (acc[0], acc[2]) and
  // (acc[1], acc[3]).
  template <typename F>
  __device__ __forceinline__ static void for_each(int row0, int col0, int m, int n,
                                                  const dbl4 (&acc)[FM][FN], F&& f) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int wm = wid % WAVES_M;
    const int wn = wid / WAVES_M;
    const int l16 = lane & 15;
    const int kk = lane >> 4;
#pragma unroll
    for (int fm = 0; fm < FM; ++fm) {
#pragma unroll
      for (int g = 0; g < FN; ++g) {
        const int j = col0 + wn * WN + 16 * g + l16;
        if (j >= n) continue;
        const dbl4 v = acc[fm][g];
        if (CPLX) {
          const int i0 = row0 / 2 + (wm * WM) / 2 + 8 * fm + kk;
          if (i0 < m) f(i0, j, v[0], v[2]);
          if (i0 + 4 < m) f(i0 + 4, j, v[1], v[3]);
        } else {
          const int p0 = row0 + wm * WM + 16 * fm + kk;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (p0 + 4 * r < m) f(p0 + 4 * r, j, v[r], 0.0);
        }
      }
    }
  }
};

template <bool CPLX>
__device__ __forceinline__ void store_work(double* w, int64_t Mp, int i, int j, double re,
                                           double im) {
  double* p = w + static_cast<int64_t>(j) * Mp + (CPLX ? 2 * i : i);
  p[0] = re;
  if (CPLX) p[1] = im;
}

}  // namespace tile
}  // namespace spla_amd
