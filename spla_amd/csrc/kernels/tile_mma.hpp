// MFMA tile routine shared by the single-product kernels (gemm_f64.hip: FP64
// and FP32, real and complex) and the persistent batched kernel (batched.hip).
// Device code only; include from HIP translation units.
//
// The scalar type S is double (v_mfma_f64_16x16x4_f64) or float
// (v_mfma_f32_16x16x4_f32). Both instructions take lane l's operands as
// A[l % 16][l / 16] and B[l / 16][l % 16]; their accumulators differ: lane l
// holds rows (l / 16) + 4 r (FP64) or 4 (l / 16) + r (FP32), r = 0..3, of
// column l % 16. A complex product puts the real and imaginary rows of a
// complex row of C into the same lane: FP64 fragment rows r and r + 8, FP32
// fragment rows 2 c and 2 c + 1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels/gemm_desc.hpp"

namespace spla_amd {
namespace tile {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef float flt2 __attribute__((ext_vector_type(2)));

template <typename S>
struct Vec;
template <>
struct Vec<double> {
  using V2 = dbl2;
  using V4 = dbl4;
};
template <>
struct Vec<float> {
  using V2 = flt2;
  using V4 = flt4;
};

constexpr int kThreads = 256;
// LDS row padding (elements; even, so that element pairs stay 16-byte aligned).
// Rows of BK = 16 doubles are 128 bytes. Measured on gfx950 with the
// bank-conflict counter on the fragment reads (ds_read_b128: 16 rows x 4
// 16-byte chunks per wave): 160-byte rows read conflict-free, 144-, 176- and
// 208-byte rows cost 4 conflict cycles per read. 160-byte rows cost 8 cycles on
// an MN-major store (eight rows, one chunk), which is rarer than the reads of
// the same tile; they fill the 160 KB of a CU exactly with two real 128 x 128
// or four real 64 x 64 workgroups. FP32 rows keep 2 floats of padding.
template <typename S, bool CPLX>
constexpr int kPadOf = sizeof(S) == 8 ? 4 : 2;

__device__ __forceinline__ double flip_sign(double x, uint32_t hiMask) {
  uint64_t b = __builtin_bit_cast(uint64_t, x);
  b ^= static_cast<uint64_t>(hiMask) << 32;
  return __builtin_bit_cast(double, b);
}
__device__ __forceinline__ float flip_sign(float x, uint32_t mask) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ mask);
}

__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma16x16x4(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// 16-byte LDS read of an element pair: the tiles are 16-byte aligned with rows
// of whole 16-byte multiples, so every pair at an even column is (ds_read_b128).
template <typename S>
__device__ __forceinline__ typename Vec<S>::V2 lds_pair(const S* p) {
  return *reinterpret_cast<const typename Vec<S>::V2*>(
      __builtin_assume_aligned(p, 2 * sizeof(S)));
}

// Stages one operand tile (ROWS stored rows x BK embedded reduction elements)
// from global memory into registers and from registers into LDS.
// UNCHECKED: k-blocks wholly inside the operands skip the per-element checks
// (measured faster on the large tiles, slower on the small ones).
template <typename S, bool CPLX, int LAYOUT, int ROWS, int BK, bool UNCHECKED>
struct TileStage {
  static constexpr int kPad = kPadOf<S, CPLX>;
  using V2 = typename Vec<S>::V2;
  static constexpr int kPairs = ROWS * BK / 2;
  static constexpr int kPerThread = kPairs / kThreads;
  static_assert(kPairs % kThreads == 0, "tile does not divide over the workgroup");
  V2 regs[kPerThread];

  __device__ __forceinline__ static void coords(int pi, int& row, int& qp) {
    if (LAYOUT == KMajor) {
      row = pi / (BK / 2);
      qp = pi % (BK / 2);
    } else {
      row = pi % ROWS;
      qp = pi / ROWS;
    }
  }

  // base: operand origin (scalars); ld: leading dimension in elements;
  // rowBase/rowLimit: stored-row range; q0/qEnd: embedded reduction range.
  // A k-block wholly inside both ranges is loaded without per-element checks.
  __device__ __forceinline__ void load(const S* __restrict__ base, int64_t ld, int rowBase,
                                       int rowLimit, int64_t q0, int64_t qEnd, int tid) {
    if (UNCHECKED && rowBase + ROWS <= rowLimit && q0 + BK <= qEnd) {
#pragma unroll
      for (int u = 0; u < kPerThread; ++u) {
        int row, qp;
        coords(tid + u * kThreads, row, qp);
        const int grow = rowBase + row;
        const int64_t q = q0 + 2 * qp;
        if (LAYOUT == KMajor) {
          const S* p = base + static_cast<int64_t>(grow) * ld * (CPLX ? 2 : 1) + q;
          regs[u] = V2{p[0], p[1]};
        } else if (CPLX) {
          const S* p = base + ((q >> 1) * ld + grow) * 2;
          regs[u] = V2{p[0], p[1]};
        } else {
          regs[u] = V2{base[q * ld + grow], base[(q + 1) * ld + grow]};
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
      S d0 = S(0), d1 = S(0);
      if (grow < rowLimit) {
        if (LAYOUT == KMajor) {
          const S* p = base + static_cast<int64_t>(grow) * ld * (CPLX ? 2 : 1) + q;
          if (q < qEnd) d0 = p[0];
          if (q + 1 < qEnd) d1 = p[1];
        } else if (CPLX) {
          if (q < qEnd) {
            const S* p = base + ((q >> 1) * ld + grow) * 2;
            d0 = p[0];
            d1 = p[1];
          }
        } else {
          if (q < qEnd) d0 = base[q * ld + grow];
          if (q + 1 < qEnd) d1 = base[(q + 1) * ld + grow];
        }
      }
      regs[u] = V2{d0, d1};
    }
  }

  __device__ __forceinline__ void store(S (*lds)[BK + kPad], int tid) const {
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      int row, qp;
      coords(tid + u * kThreads, row, qp);
      *reinterpret_cast<V2*>(&lds[row][2 * qp]) = regs[u];
    }
  }
};

// Operands of one tile product (element rows m / columns n of C).
template <typename S>
struct TileOperands {
  const S* A;
  const S* B;
  int64_t lda, ldb;
  int m, n;
  int conjA, conjB;
};

// LDS of one tile: two buffers of AR stored A rows and BN B rows of BK elements.
// Rows are whole 16-byte multiples; the 16-byte alignment lets the fragment
// reads compile to ds_read_b128.
template <typename S, bool CPLX, int AR, int BN, int BK>
struct TileSmem {
  alignas(16) S lA[2][AR][BK + kPadOf<S, CPLX>];
  alignas(16) S lB[2][BN][BK + kPadOf<S, CPLX>];
};

// One BM x BN tile of the embedded product over the embedded reduction range
// [kBegin, kEnd), accumulated in registers. Shared by the single-product
// kernels and the batched kernels. On return every thread has passed the final
// barrier, so the LDS buffers may be reused at once (reads that one schedule
// issues after that barrier only fill registers that are discarded).
// LEAN: for callers that keep more state live across tiles (the persistent
// batched kernel), which only take the schedules with one fragment set and one
// barrier per k-block.
template <typename S, bool CPLX, int LA, int LB, int WM, int WN, int WAVES_M, int BK,
          bool LEAN = false>
struct TileMma {
  using V2 = typename Vec<S>::V2;
  using Acc = typename Vec<S>::V4;
  static constexpr int WAVES_N = 4 / WAVES_M;
  static constexpr int BM = WM * WAVES_M;
  static constexpr int BN = WN * WAVES_N;
  static constexpr int AR = CPLX ? BM / 2 : BM;
  static constexpr int FM = WM / 16;
  static constexpr int FN = WN / 16;
  static_assert(BK % 8 == 0, "k-block must hold whole 8-element sub-blocks");
  static constexpr bool kF64 = sizeof(S) == 8;

  // Complex fragment row r (= lane % 16 of the A operand): its complex row
  // inside the fragment and whether it is an imaginary row.
  __device__ __forceinline__ static int cplx_row(int r) { return kF64 ? (r & 7) : (r >> 1); }
  __device__ __forceinline__ static bool imag_row(int r) { return kF64 ? r >= 8 : (r & 1) != 0; }

  static constexpr int kPad = kPadOf<S, CPLX>;
  using Smem = TileSmem<S, CPLX, AR, BN, BK>;

  __device__ __forceinline__ static void run(const TileOperands<S>& op, int row0, int col0,
                                             int64_t kBegin, int64_t kEnd, Smem& sm,
                                             Acc (&acc)[FM][FN]) {
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
    // Fragment row r = lane % 16 is an imaginary row when imag_row(r).
    bool swapPair = false;
    uint32_t sgn0 = 0, sgn1 = 0;
    if (CPLX) {
      swapPair = imag_row(l16);
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
      for (int g = 0; g < FN; ++g) acc[f][g] = Acc{S(0), S(0), S(0), S(0)};

    constexpr bool kUnchecked = WM * WN >= 64 * 64;
    TileStage<S, CPLX, LA, AR, BK, kUnchecked> stA;
    TileStage<S, CPLX, LB, BN, BK, kUnchecked> stB;

    // Large tiles: the next 8-element sub-block's fragments are read from LDS while
    // the current one's MFMAs run (two register sets; 2.6 % faster on the complex
    // FP64 A^H B tiles with one barrier per k-block, but 6 % slower on FP32 and 0.5 %
    // on the small ring-step tiles). With the split-barrier schedule below the 64 x 64
    // wave tiles gain more: FP64 MN-major A (A B) takes it on complex tiles only
    // (pzgemm_sbs 100000 x 4096 x 4096 63.6 -> 66.6 TFLOP/s; the real ones spill);
    // FP32 in every layout but real MN-major B, at two waves per SIMD (2048 x 2048 x
    // 1e5: complex A^H B 110.7 -> 120.3 TFLOP/s, real A^T B 111.6 -> 121.5, real A B
    // 103.4 -> 118.5; three waves per SIMD, 168 VGPRs: 109.6). Real tiles with
    // MN-major B read just in time: with the second set they spilled 132 VGPRs in
    // FP64 (2048 x 2048 x 1e5 A^T B^T 44.7 TFLOP/s, 59.6 without).
    constexpr bool kPrefetch =
        (kF64 && WM * WN >= 64 * 64 && (CPLX || LB == KMajor) &&
         (LA == KMajor || (!LEAN && CPLX && WM * WN == 64 * 64))) ||
        (!kF64 && !LEAN && BK == 16 && WM * WN == 64 * 64 && (CPLX || LB == KMajor));
    if constexpr (kPrefetch) {
      static_assert((BK / 8) % 2 == 0, "sub-blocks alternate between two register sets");
      S a0[2][FM], a1[2][FM];
      V2 b[2][FN];
      auto read_frags = [&](int set, int buf, int kb8) {
#pragma unroll
        for (int f = 0; f < FM; ++f) {
          const int sr = CPLX ? (wm * WM) / 2 + 8 * f + cplx_row(l16) : wm * WM + 16 * f + l16;
          const V2 v = lds_pair(&sm.lA[buf][sr][kb8 * 8 + 2 * kk]);
          if (CPLX) {
            const S x = swapPair ? v.y : v.x;
            const S y = swapPair ? v.x : v.y;
            a0[set][f] = flip_sign(x, sgn0);
            a1[set][f] = flip_sign(y, sgn1);
          } else {
            a0[set][f] = v.x;
            a1[set][f] = v.y;
          }
        }
#pragma unroll
        for (int g = 0; g < FN; ++g) {
          const int j = wn * WN + 16 * g + l16;
          b[set][g] = lds_pair(&sm.lB[buf][j][kb8 * 8 + 2 * kk]);
        }
      };

      // half = 0: the even reduction elements of a sub-block, 1: the odd ones
      auto mfma_half = [&](int set, int half) {
#pragma unroll
        for (int f = 0; f < FM; ++f)
#pragma unroll
          for (int g = 0; g < FN; ++g)
            acc[f][g] = half ? mfma16x16x4(a1[set][f], b[set][g].y, acc[f][g])
                             : mfma16x16x4(a0[set][f], b[set][g].x, acc[f][g]);
      };

      if (nk > 0) {
        stA.load(op.A, op.lda, arow0, op.m, kBegin, kEnd, tid);
        stB.load(op.B, op.ldb, col0, op.n, kBegin, kEnd, tid);
        stA.store(sm.lA[0], tid);
        stB.store(sm.lB[0], tid);
      }
      __syncthreads();
      if (nk > 0) read_frags(0, 0, 0);

      // 64 x 64 wave tiles (two waves per SIMD) except real MN-major B, whose staging
      // leaves no registers for it; the 128 x 64 tiles spill with it.
      constexpr bool kSplitBarrier =
          !LEAN && BK == 16 && WM * WN == 64 * 64 && (CPLX || LB == KMajor);
      if constexpr (kSplitBarrier) {
        // A k-block's 4 FM FN MFMAs in quarters: (sub-block 0, even elements) with the
        // reads of sub-block 1; the staged stores into the other buffer; (sub-block 0,
        // odd), which covers the stores' completion; the barrier; the next k-block's
        // sub-block 0 reads; both quarters of sub-block 1, which cover those reads.
        // Against one barrier after all the MFMAs of a k-block, with the stores and
        // reads on either side of it: 2048 x 2048 x 2e5 complex A^H B 109.8 -> 102.7 ms.
        if (nk > 1) {
          stA.load(op.A, op.lda, arow0, op.m, kBegin + BK, kEnd, tid);
          stB.load(op.B, op.ldb, col0, op.n, kBegin + BK, kEnd, tid);
        }
        // k-block it + 2 into the staging registers once they are stored. Real tiles
        // (eight staged pairs per thread) issue these loads right after the stores;
        // complex tiles (six) after the barrier and the next fragment reads, which
        // measured 1-2 % faster for each.
        // The stores and the reads after the barrier are unconditional (branches here
        // make the wait counts conservative): in the last k-block they store the
        // staged registers' last k-block once more into the buffer that is no longer
        // read, and read them back into registers that are not used (reads issued
        // after the final barrier whose values are discarded).
        for (int it = 0; it < nk; ++it) {
          const int cur = it & 1;
          read_frags(1, cur, 1);
          mfma_half(0, 0);
          __builtin_amdgcn_sched_barrier(0);
          stA.store(sm.lA[cur ^ 1], tid);
          stB.store(sm.lB[cur ^ 1], tid);
          if constexpr (!CPLX) {
            if (it + 2 < nk) {
              const int64_t q0 = kBegin + static_cast<int64_t>(it + 2) * BK;
              stA.load(op.A, op.lda, arow0, op.m, q0, kEnd, tid);
              stB.load(op.B, op.ldb, col0, op.n, q0, kEnd, tid);
            }
          }
          mfma_half(0, 1);
          __builtin_amdgcn_sched_barrier(0);  // these MFMAs cover the stores' completion
          __syncthreads();
          read_frags(0, cur ^ 1, 0);
          if constexpr (CPLX) {
            if (it + 2 < nk) {
              const int64_t q0 = kBegin + static_cast<int64_t>(it + 2) * BK;
              stA.load(op.A, op.lda, arow0, op.m, q0, kEnd, tid);
              stB.load(op.B, op.ldb, col0, op.n, q0, kEnd, tid);
            }
          }
          mfma_half(1, 0);
          mfma_half(1, 1);
        }
        return;
      }

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
          const int s = kb8 & 1;
          if (kb8 + 1 < BK / 8) read_frags(s ^ 1, cur, kb8 + 1);
#pragma unroll
          for (int f = 0; f < FM; ++f)
#pragma unroll
            for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a0[s][f], b[s][g].x, acc[f][g]);
#pragma unroll
          for (int f = 0; f < FM; ++f)
#pragma unroll
            for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a1[s][f], b[s][g].y, acc[f][g]);
        }

        if (more) {
          stA.store(sm.lA[cur ^ 1], tid);
          stB.store(sm.lB[cur ^ 1], tid);
        }
        __syncthreads();
        if (more) read_frags(0, cur ^ 1, 0);  // the next k-block's first sub-block
      }
    } else if constexpr (WM * WN < 64 * 64 && (CPLX || LA != MNMajor || LB != MNMajor)) {
      // Small tiles (small m, n; HBM-bound with a long reduction): the global loads
      // run two k-blocks ahead through two register sets, so each workgroup keeps
      // twice the bytes in flight; fragments are read just in time. The real
      // MN-major x MN-major variant stays one ahead: in FP64 the second set takes it
      // past 128 VGPRs, one resident workgroup fewer per CU, and it ran 7 % slower.
      TileStage<S, CPLX, LA, AR, BK, kUnchecked> stA2;
      TileStage<S, CPLX, LB, BN, BK, kUnchecked> stB2;
      if (nk > 0) {
        stA.load(op.A, op.lda, arow0, op.m, kBegin, kEnd, tid);
        stB.load(op.B, op.ldb, col0, op.n, kBegin, kEnd, tid);
        stA.store(sm.lA[0], tid);
        stB.store(sm.lB[0], tid);
      }
      if (nk > 1) {
        stA2.load(op.A, op.lda, arow0, op.m, kBegin + BK, kEnd, tid);
        stB2.load(op.B, op.ldb, col0, op.n, kBegin + BK, kEnd, tid);
      }
      __syncthreads();
      if constexpr (kF64 && BK == 16 && !LEAN) {
        // The split-barrier schedule of the large tiles with the loads two ahead:
        // fragments in two register sets (sub-block 1's read during sub-block 0's
        // MFMAs, the next k-block's sub-block 0 right after the barrier), the staged
        // stores between the two halves of sub-block 0's MFMAs, sub-block 1's after
        // the barrier. Stores and post-barrier reads are unconditional as there; a
        // staging set never loaded (one k-block) stores zeros.
        S fa0[2][FM], fa1[2][FM];
        V2 fb[2][FN];
        auto rd = [&](int set, int buf, int kb8) {
#pragma unroll
          for (int f = 0; f < FM; ++f) {
            const int sr = CPLX ? (wm * WM) / 2 + 8 * f + cplx_row(l16) : wm * WM + 16 * f + l16;
            const V2 v = lds_pair(&sm.lA[buf][sr][kb8 * 8 + 2 * kk]);
            if (CPLX) {
              const S x = swapPair ? v.y : v.x;
              const S y = swapPair ? v.x : v.y;
              fa0[set][f] = flip_sign(x, sgn0);
              fa1[set][f] = flip_sign(y, sgn1);
            } else {
              fa0[set][f] = v.x;
              fa1[set][f] = v.y;
            }
          }
#pragma unroll
          for (int g = 0; g < FN; ++g) {
            const int j = wn * WN + 16 * g + l16;
            fb[set][g] = lds_pair(&sm.lB[buf][j][kb8 * 8 + 2 * kk]);
          }
        };
        auto mm = [&](int set, int half) {
#pragma unroll
          for (int f = 0; f < FM; ++f)
#pragma unroll
            for (int g = 0; g < FN; ++g)
              acc[f][g] = half ? mfma16x16x4(fa1[set][f], fb[set][g].y, acc[f][g])
                               : mfma16x16x4(fa0[set][f], fb[set][g].x, acc[f][g]);
        };
        if (nk == 1) {
#pragma unroll
          for (int u = 0; u < decltype(stA2)::kPerThread; ++u) stA2.regs[u] = V2{S(0), S(0)};
#pragma unroll
          for (int u = 0; u < decltype(stB2)::kPerThread; ++u) stB2.regs[u] = V2{S(0), S(0)};
        }
        if (nk > 0) rd(0, 0, 0);
        auto kblock2 = [&](int it, auto& la, auto& lb, const auto& sa, const auto& sb) {
          const int cur = it & 1;
          if (it + 2 < nk) {
            const int64_t q0 = kBegin + static_cast<int64_t>(it + 2) * BK;
            la.load(op.A, op.lda, arow0, op.m, q0, kEnd, tid);
            lb.load(op.B, op.ldb, col0, op.n, q0, kEnd, tid);
          }
          rd(1, cur, 1);
          mm(0, 0);
          __builtin_amdgcn_sched_barrier(0);
          sa.store(sm.lA[cur ^ 1], tid);
          sb.store(sm.lB[cur ^ 1], tid);
          mm(0, 1);
          __builtin_amdgcn_sched_barrier(0);
          __syncthreads();
          rd(0, cur ^ 1, 0);
          mm(1, 0);
          mm(1, 1);
        };
        for (int it = 0; it < nk; it += 2) {
          kblock2(it, stA, stB, stA2, stB2);
          if (it + 1 < nk) kblock2(it + 1, stA2, stB2, stA, stB);
        }
        return;
      }
      // one k-block: (la, lb) receive k-block it + 2, (sa, sb) hold it + 1
      auto kblock = [&](int it, auto& la, auto& lb, const auto& sa, const auto& sb) {
        const int cur = it & 1;
        if (it + 2 < nk) {
          const int64_t q0 = kBegin + static_cast<int64_t>(it + 2) * BK;
          la.load(op.A, op.lda, arow0, op.m, q0, kEnd, tid);
          lb.load(op.B, op.ldb, col0, op.n, q0, kEnd, tid);
        }
#pragma unroll
        for (int kb8 = 0; kb8 < BK / 8; ++kb8) {
          S a0[FM], a1[FM];
          V2 b[FN];
#pragma unroll
          for (int f = 0; f < FM; ++f) {
            const int sr = CPLX ? (wm * WM) / 2 + 8 * f + cplx_row(l16) : wm * WM + 16 * f + l16;
            const V2 v = lds_pair(&sm.lA[cur][sr][kb8 * 8 + 2 * kk]);
            if (CPLX) {
              const S x = swapPair ? v.y : v.x;
              const S y = swapPair ? v.x : v.y;
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
            b[g] = lds_pair(&sm.lB[cur][j][kb8 * 8 + 2 * kk]);
          }
#pragma unroll
          for (int f = 0; f < FM; ++f)
#pragma unroll
            for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a0[f], b[g].x, acc[f][g]);
#pragma unroll
          for (int f = 0; f < FM; ++f)
#pragma unroll
            for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a1[f], b[g].y, acc[f][g]);
        }
        if (it + 1 < nk) {
          sa.store(sm.lA[cur ^ 1], tid);
          sb.store(sm.lB[cur ^ 1], tid);
        }
        __syncthreads();
      };
      for (int it = 0; it < nk; it += 2) {
        kblock(it, stA, stB, stA2, stB2);
        if (it + 1 < nk) kblock(it + 1, stA2, stB2, stA, stB);
      }
    } else {
      // each sub-block's fragments are read right before its MFMAs
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
          S a0[FM], a1[FM];
          V2 b[FN];
#pragma unroll
          for (int f = 0; f < FM; ++f) {
            const int sr = CPLX ? (wm * WM) / 2 + 8 * f + cplx_row(l16) : wm * WM + 16 * f + l16;
            const V2 v = lds_pair(&sm.lA[cur][sr][kb8 * 8 + 2 * kk]);
            if (CPLX) {
              const S x = swapPair ? v.y : v.x;
              const S y = swapPair ? v.x : v.y;
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
            b[g] = lds_pair(&sm.lB[cur][j][kb8 * 8 + 2 * kk]);
          }
#pragma unroll
          for (int f = 0; f < FM; ++f)
#pragma unroll
            for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a0[f], b[g].x, acc[f][g]);
#pragma unroll
          for (int f = 0; f < FM; ++f)
#pragma unroll
            for (int g = 0; g < FN; ++g) acc[f][g] = mfma16x16x4(a1[f], b[g].y, acc[f][g]);
        }

        if (more) {
          stA.store(sm.lA[cur ^ 1], tid);
          stB.store(sm.lB[cur ^ 1], tid);
        }
        __syncthreads();
      }
    }
  }

  // Calls f(i, j, re, im) for every element of the tile inside m x n.
  // Accumulators of lane (kk = lane / 16) of column lane % 16: FP64 rows
  // kk + 4 r, i.e. (re, im) of complex rows kk and kk + 4 in (acc[0], acc[2])
  // and (acc[1], acc[3]); FP32 rows 4 kk + r, i.e. (re, im) of complex rows
  // 2 kk and 2 kk + 1 in (acc[0], acc[1]) and (acc[2], acc[3]).
  template <typename F>
  __device__ __forceinline__ static void for_each(int row0, int col0, int m, int n,
                                                  const Acc (&acc)[FM][FN], F&& f) {
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
        const Acc v = acc[fm][g];
        if (CPLX) {
          const int base = row0 / 2 + (wm * WM) / 2 + 8 * fm;
          if (kF64) {
            if (base + kk < m) f(base + kk, j, v[0], v[2]);
            if (base + kk + 4 < m) f(base + kk + 4, j, v[1], v[3]);
          } else {
            if (base + 2 * kk < m) f(base + 2 * kk, j, v[0], v[1]);
            if (base + 2 * kk + 1 < m) f(base + 2 * kk + 1, j, v[2], v[3]);
          }
        } else {
          const int p0 = row0 + wm * WM + 16 * fm + (kF64 ? kk : 4 * kk);
          const int step = kF64 ? 4 : 1;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (p0 + step * r < m) f(p0 + step * r, j, v[r], S(0));
        }
      }
    }
  }
};

template <typename S, bool CPLX>
__device__ __forceinline__ void store_work(S* w, int64_t Mp, int i, int j, S re, S im) {
  S* p = w + static_cast<int64_t>(j) * Mp + (CPLX ? 2 * i : i);
  p[0] = re;
  if (CPLX) p[1] = im;
}

}  // namespace tile
}  // namespace spla_amd
