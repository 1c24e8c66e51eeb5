// Split-precision (Ozaki scheme II) kernels for gfx950: FP64-accurate products
// on the INT8 matrix cores (v_mfma_i32_16x16x64_i8, or v_mfma_i32_32x32x32_i8
// with SPLA_AMD_I8_MFMA=32). See ozaki_desc.hpp for the arithmetic.
//
//   stats_*      per-row largest |component|, sum of squares, non-finite flag
//                (HBM bound: reads each operand once);
//   residue_*    scaled integer rows -> symmetric residues modulo every m_t
//                (residue_arith.hpp), written as K-major int8 planes padded to
//                256 bytes (HBM bound: reads the operand once, writes
//                N*products bytes per element);
//   i8_gemm      batched C_g = R_A,g R_B,g^T over all moduli and complex
//                sub-products in one launch, INT32 accumulation; splits of at
//                most 2^17 - 256 reduction bytes keep the sums below 2^31
//                (|residue| <= 128). One split: the epilogue reduces modulo
//                m_t and stores biased bytes r + 128; several splits: raw
//                int32 sums;
//   crt          per four output rows: modular recombination of the complex
//                sub-products, Chinese remainder reconstruction exact in FP64
//                limb sums and 192-bit integer assembly (crt_arith.hpp),
//                conversion to FP64, scaling by 2^-(ea_i + eb_j) and the
//                alpha / beta update.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>

#include "kernels/crt_arith.hpp"
#include "kernels/launch.hpp"
#include "kernels/ozaki_desc.hpp"
#include "kernels/residue_arith.hpp"

namespace spla_amd {
namespace {

using residue::mix;
using residue::ModConst;
using residue::Pieces;
using residue::pieces14;
using residue::residue_biased;

typedef int int4v __attribute__((ext_vector_type(4)));
typedef int int16v __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// statistics

template <bool CPLX>
__global__ void __launch_bounds__(256) stats_kmajor_kernel(const double* __restrict__ X,
                                                           int64_t ld, int rows, int64_t Kp,
                                                           int64_t perSplit, double* pmax,
                                                           double* psum) {
  const int r = blockIdx.x;
  const int s = blockIdx.y;
  const int64_t q0 = static_cast<int64_t>(s) * perSplit;
  const int64_t q1 = min(Kp, q0 + perSplit);
  const double* row = X + static_cast<int64_t>(r) * ld * (CPLX ? 2 : 1);
  double mx = 0.0, sq = 0.0;
  int bad = 0;
  auto take = [&](double x) {
    bad |= !isfinite(x);
    mx = fmax(mx, fabs(x));
    sq += x * x;
  };
  // 16-byte non-temporal loads (the operand is read once more, by the residue
  // pass, long after it has left the caches): one leading element when the range
  // starts at an odd double, pairs, one trailing element
  const double* p = row + q0;
  int64_t cnt = q1 - q0;
  if (cnt > 0 && (reinterpret_cast<uintptr_t>(p) & 15) != 0) {
    if (threadIdx.x == 0) take(p[0]);
    ++p;
    --cnt;
  }
  typedef double d2v __attribute__((ext_vector_type(2)));
  const d2v* p2 = reinterpret_cast<const d2v*>(p);
  const int64_t pairs = cnt / 2;
#pragma unroll 4
  for (int64_t j = threadIdx.x; j < pairs; j += 256) {
    const d2v v = __builtin_nontemporal_load(p2 + j);
    take(v.x);
    take(v.y);
  }
  if ((cnt & 1) != 0 && threadIdx.x == 0) take(p[cnt - 1]);
  __shared__ double smx[256], ssq[256];
  __shared__ int sbad[256];
  smx[threadIdx.x] = mx;
  ssq[threadIdx.x] = sq;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + w]);
      ssq[threadIdx.x] += ssq[threadIdx.x + w];
      sbad[threadIdx.x] |= sbad[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[static_cast<int64_t>(s) * rows + r] = sbad[0] ? __builtin_nan("") : smx[0];
    psum[static_cast<int64_t>(s) * rows + r] = ssq[0];
  }
}

// Rows contiguous in memory: 64 rows per block (one per lane), 4 waves split the
// reduction index.
template <bool CPLX>
__global__ void __launch_bounds__(256) stats_mnmajor_kernel(const double* __restrict__ X,
                                                            int64_t ld, int rows, int64_t k,
                                                            int64_t perSplit, double* pmax,
                                                            double* psum) {
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + tx;
  const int s = blockIdx.y;
  const int64_t l0 = static_cast<int64_t>(s) * perSplit;
  const int64_t l1 = min(k, l0 + perSplit);
  double mx = 0.0, sq = 0.0;
  int bad = 0;
  if (r < rows) {
    for (int64_t l = l0 + ty; l < l1; l += 4) {
      if (CPLX) {
        const double* p = X + (l * ld + r) * 2;
        const double a = p[0], b = p[1];
        bad |= !isfinite(a) | !isfinite(b);
        mx = fmax(mx, fmax(fabs(a), fabs(b)));
        sq += a * a + b * b;
      } else {
        const double a = X[l * ld + r];
        bad |= !isfinite(a);
        mx = fmax(mx, fabs(a));
        sq += a * a;
      }
    }
  }
  __shared__ double smx[4][64], ssq[4][64];
  __shared__ int sbad[4][64];
  smx[ty][tx] = mx;
  ssq[ty][tx] = sq;
  sbad[ty][tx] = bad;
  __syncthreads();
  if (ty == 0 && r < rows) {
    for (int w = 1; w < 4; ++w) {
      mx = fmax(mx, smx[w][tx]);
      sq += ssq[w][tx];
      bad |= sbad[w][tx];
    }
    pmax[static_cast<int64_t>(s) * rows + r] = bad ? __builtin_nan("") : mx;
    psum[static_cast<int64_t>(s) * rows + r] = sq;
  }
}

__global__ void __launch_bounds__(256) stats_finalize_kernel(const double* pmax,
                                                             const double* psum, int rows,
                                                             int splits, double* omax,
                                                             double* osum) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  double mx = 0.0, sq = 0.0;
  int bad = 0;
  for (int s = 0; s < splits; ++s) {
    const double v = pmax[static_cast<int64_t>(s) * rows + r];
    bad |= isnan(v);
    mx = fmax(mx, v);
    sq += psum[static_cast<int64_t>(s) * rows + r];
  }
  omax[r] = bad ? __builtin_nan("") : mx;
  osum[r] = sq;
}

// ---------------------------------------------------------------------------
// residues

// Biased residue rb = r + 128 in [0, 255] (exact float) into byte j of w.
__device__ __forceinline__ uint32_t put_byte(float rb, int j, uint32_t w) {
  return __builtin_amdgcn_cvt_pk_u8_f32(rb, j, w);
}

// Four biased bytes -> four int8 residues (r + 128 with the top bit flipped is r).
constexpr uint32_t kByteSigns = 0x80808080u;

// The three planes of one element and modulus: Re (from the biased pieces,
// p0 + 128), Im (unbiased pieces) and (Re + Im) mod m, each as a biased byte.
struct Biased3 {
  float re, im, sum;
};
__device__ __forceinline__ Biased3 residues3(const Pieces& reb, const Pieces& im,
                                             const ModConst& kc) {
  const float rre = residue_biased(mix(reb, reb.p0, kc), kc);
  const float rim = residue::residue(mix(im, im.p0, kc), kc);
  return {rre, rim + 128.0f, residue_biased(rre + rim, kc)};
}

// Reduction index contiguous: a block converts 1024 consecutive elements of a
// row. The row segment is read with whole-wave contiguous 16-byte loads into
// LDS, then every thread takes 4 consecutive elements and writes one 32-bit
// word per residue plane (a wave writes 256 contiguous bytes per plane).
constexpr int kResElems = 1024;  // elements per block
__device__ __forceinline__ int res_slot(int e) { return e + (e >> 3); }  // conflict-free reads

template <bool CPLX>
__global__ void __launch_bounds__(256) residue_kmajor_kernel(ResidueDesc d) {
  constexpr int C = CPLX ? 2 : 1;  // doubles per element
  // 16-byte slots: one complex element, or two real elements
  constexpr int kSlots = kResElems * C / 2;
  __shared__ __attribute__((aligned(16))) double2 lds[kSlots + kSlots / 8];
  const int row = blockIdx.y;
  const int64_t e0 = static_cast<int64_t>(blockIdx.x) * kResElems;
  const int tid = threadIdx.x;
  const double* src = d.X + static_cast<int64_t>(row) * d.ld * C + e0 * C;
  const int64_t avail = (d.k - e0) * C;  // doubles of the row left from e0 (may be <= 0)
  const bool aligned = (reinterpret_cast<uintptr_t>(src) & 15) == 0;  // uniform per block
#pragma unroll
  for (int j = 0; j < kSlots / 256; ++j) {
    const int slot = tid + 256 * j;
    double2 v = {0.0, 0.0};
    if (aligned && 2 * slot + 1 < avail) {
      // streamed once: non-temporal loads and stores (residue pass 1.5 % faster)
      typedef double d2v __attribute__((ext_vector_type(2)));
      const d2v t = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(src + 2 * slot));
      v = double2{t.x, t.y};
    } else {
      if (2 * slot < avail) v.x = src[2 * slot];
      if (2 * slot + 1 < avail) v.y = src[2 * slot + 1];
    }
    lds[res_slot(slot)] = v;
  }
  __syncthreads();
  const int64_t l0 = e0 + 4 * tid;  // first of this thread's 4 elements
  if (l0 >= d.kpad) return;
  const int e = d.expo[row];
  Pieces re[4], im[4];  // re carries the +128 of the biased form in p0
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (CPLX) {
      const double2 v = lds[res_slot(4 * tid + j)];
      re[j] = pieces14(rint(ldexp(v.x, e)));
      im[j] = pieces14(d.imagSign * rint(ldexp(v.y, e)));
    } else {
      const double2 v = lds[res_slot(2 * tid + (j >> 1))];
      re[j] = pieces14(rint(ldexp((j & 1) ? v.y : v.x, e)));
      im[j] = Pieces{0.f, 0.f, 0.f, 0.f};
    }
    re[j].p0 += 128.0f;
  }
  constexpr int P = CPLX ? 3 : 1;
  int8_t* base = d.R + static_cast<int64_t>(row) * d.kpad + l0;
  for (int t = 0; t < d.ms.count; ++t) {
    const ModConst kc = d.ms.kc[t];
    uint32_t w0 = 0, w1 = 0, w2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (CPLX) {
        const Biased3 r = residues3(re[j], im[j], kc);
        w0 = put_byte(r.re, j, w0);
        w1 = put_byte(r.im, j, w1);
        w2 = put_byte(r.sum, j, w2);
      } else {
        w0 = put_byte(residue_biased(mix(re[j], re[j].p0, kc), kc), j, w0);
      }
    }
    int8_t* p = base + static_cast<int64_t>(t * P) * d.plane;
    __builtin_nontemporal_store(w0 ^ kByteSigns, reinterpret_cast<uint32_t*>(p));
    if (CPLX) {
      __builtin_nontemporal_store(w1 ^ kByteSigns, reinterpret_cast<uint32_t*>(p + d.plane));
      __builtin_nontemporal_store(w2 ^ kByteSigns, reinterpret_cast<uint32_t*>(p + 2 * d.plane));
    }
  }
}

// Rows contiguous: each lane converts 16 consecutive reduction indices of one
// row, which is exactly one 16-byte chunk of that row in every K-major plane, so
// the transposition happens in registers (no LDS, no barrier). Eight lanes cover
// the 128 reduction indices of a row, a wave eight rows: every load instruction
// reads eight 128-byte row segments (one per reduction index) and every store
// writes eight whole 128-byte plane rows. A block of 4 waves converts a
// 32-row x 128-index tile.
constexpr int kResMnK = 128;    // reduction indices per block
constexpr int kResMnRows = 32;  // rows per block
template <bool CPLX>
__global__ void __launch_bounds__(256) residue_mnmajor_kernel(ResidueDesc d) {
  constexpr int P = CPLX ? 3 : 1;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.y * kResMnRows + 8 * (threadIdx.x >> 6) + (lane >> 3);
  const int64_t k0 = static_cast<int64_t>(blockIdx.x) * kResMnK + 16 * (lane & 7);
  const bool live = row < d.rows;
  const int e = live ? d.expo[row] : 0;
  Pieces re[16], im[16];  // re carries the +128 of the biased form in p0
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int64_t l = k0 + j;
    double vr = 0.0, vi = 0.0;
    if (live && l < d.k) {
      if (CPLX) {
        const double* p = d.X + (l * d.ld + row) * 2;
        vr = rint(ldexp(p[0], e));
        vi = d.imagSign * rint(ldexp(p[1], e));
      } else {
        vr = rint(ldexp(d.X[l * d.ld + row], e));
      }
    }
    re[j] = pieces14(vr);
    re[j].p0 += 128.0f;
    im[j] = pieces14(vi);
  }
  if (!live) return;
  int8_t* base = d.R + static_cast<int64_t>(row) * d.kpad + k0;
  for (int t = 0; t < d.ms.count; ++t) {
    const ModConst kc = d.ms.kc[t];
    int4v w[P];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t w0 = 0, w1 = 0, w2 = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = 4 * q + b;
        if (CPLX) {
          const Biased3 r = residues3(re[j], im[j], kc);
          w0 = put_byte(r.re, b, w0);
          w1 = put_byte(r.im, b, w1);
          w2 = put_byte(r.sum, b, w2);
        } else {
          w0 = put_byte(residue_biased(mix(re[j], re[j].p0, kc), kc), b, w0);
        }
      }
      w[0][q] = static_cast<int>(w0 ^ kByteSigns);
      if (CPLX) {
        w[1 % P][q] = static_cast<int>(w1 ^ kByteSigns);
        w[2 % P][q] = static_cast<int>(w2 ^ kByteSigns);
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
      *reinterpret_cast<int4v*>(base + static_cast<int64_t>(t * P + p) * d.plane) = w[p];
  }
}

// ---------------------------------------------------------------------------
// batched INT8 GEMM

__device__ __forceinline__ bool tile_active(int mode, int64_t rowOff, int64_t colOff, int64_t r,
                                            int64_t c, int64_t nr, int64_t nc) {
  // SPLA block_is_active (src/block_generation/block.hpp:58-66) on a tile.
  if (mode == 1) return c + nc + colOff > r + rowOff;
  if (mode == 2) return r + nr + rowOff > c + colOff;
  return true;
}

// Residue planes hold whole tiles: rows are padded to a multiple of the tile
// height (256) so that every load of a tile is in bounds without masks; padded
// rows only feed output rows / columns that are never stored.
//
// Pipeline per k-block (BK = 128 bytes): global loads run about two k-blocks
// ahead through two register sets, LDS is double buffered (one barrier per
// k-block), and every memory instruction of a step (global loads, LDS writes of
// the next k-block, LDS operand reads of the next 32-byte step) is interleaved
// with its 64 MFMAs.
// Accumulators of one MFMA: 32 x 32 (16 registers) or 16 x 16 (4 registers).
template <bool M16>
struct I8Acc {
  typedef int16v type;
  static constexpr int regs = 16;
};
template <>
struct I8Acc<true> {
  typedef int4v type;
  static constexpr int regs = 4;
};

// v_mfma_i32_16x16x64_i8 with the accumulator tied to its AGPRs. The compiler
// has no tied (accumulate-in-place) form of 16 x 16 MFMAs, and with several
// updates of every accumulator per loop iteration its register allocator
// rotates the 64 accumulators of a 128 x 128 wave tile through the register
// file (hundreds of AGPR moves per k-block). Issued from inline assembly the
// accumulator stays in place; as a volatile statement it also keeps its place
// among the memory instructions, so the issue order is the source order.
// Hazards the compiler no longer sees: the operands come from LDS reads (the
// compiler waits for them), an accumulator is updated again only 63 MFMAs
// later, and the accumulators are read back only after explicit wait states.
__device__ __forceinline__ void i8_mfma16(int4v& c, const int4v& a, const int4v& b) {
  asm volatile("v_mfma_i32_16x16x64_i8 %0, %1, %2, %0" : "+a"(c) : "v"(a), "v"(b));
}

// M16 selects v_mfma_i32_16x16x64_i8 over v_mfma_i32_32x32x32_i8. Both do the
// same work per matrix-core cycle, but on random residue bytes a register-only
// loop of the 16 x 16 x 64 form sustains ~4070 TOP/s against ~3430 for the
// 32 x 32 x 32 form under the same power limit (constant bytes: ~4950 both):
// it touches half the accumulator bits per multiply-add.
template <bool OUT8, int WM, int WN, bool M16>
__global__ void __launch_bounds__(256, 1) i8_gemm_kernel(I8GemmDesc d) {
  constexpr int BM = 2 * WM;
  constexpr int BN = 2 * WN;
  constexpr int BK = 128;
  // LDS row stride: the 32 x 32 fragments (32 rows x 2 chunks of 16 bytes per
  // wave read) are conflict-free at 144 bytes, the 16 x 16 ones (16 rows x 4
  // chunks) only at 160 bytes (bank-conflict counter: 144 bytes cost ~2.7
  // cycles per LDS instruction); 2 x 512 rows x 160 bytes is the whole 160 KiB
  constexpr int kStride = BK + (M16 ? 32 : 16);
  constexpr int FM = WM / (M16 ? 16 : 32);
  constexpr int FN = WN / (M16 ? 16 : 32);
  typedef typename I8Acc<M16>::type AccV;
  constexpr int kAccRegs = I8Acc<M16>::regs;
  constexpr int CA = BM / 32;  // 16-byte chunks per thread and k-block (32 rows per pass)
  constexpr int CB = BN / 32;
  __shared__ __attribute__((aligned(16))) int8_t lA[2][BM][kStride];
  __shared__ __attribute__((aligned(16))) int8_t lB[2][BN][kStride];

  // block -> (split, product, tile); consecutive blocks land on one XCD after the remap
  const int G = gridDim.x;
  const int bid = blockIdx.x;
  int lin = bid;
  if ((G & 7) == 0) lin = (bid & 7) * (G >> 3) + (bid >> 3);
  const int tiles = d.tilesM * d.tilesN;
  const int s = lin / (d.G * tiles);
  int rest = lin - s * d.G * tiles;
  const int g = rest / tiles;
  const int t = rest - g * tiles;
  const int group = t / (8 * d.tilesN);
  const int firstM = group * 8;
  const int gsize = min(d.tilesM - firstM, 8);
  const int inGroup = t - group * 8 * d.tilesN;
  const int tm = firstM + inGroup % gsize;
  const int tn = inGroup / gsize;
  const int row0 = tm * BM;
  const int col0 = tn * BN;
  if (d.fillMode != 0 &&
      !tile_active(d.fillMode, d.fillRowOff, d.fillColOff, row0, col0, min(BM, d.m - row0),
                   min(BN, d.n - col0)))
    return;

  const int64_t kBegin = static_cast<int64_t>(s) * d.kSplit;
  const int64_t kEnd = min(d.kpad, kBegin + d.kSplit);
  const int nk = static_cast<int>((kEnd - kBegin) / BK);
  const int64_t kpad = d.kpad;
  // tile origins (uniform) and the thread's offset inside a 32-row pass
  const int8_t* At = d.RA + static_cast<int64_t>(g) * d.planeA + static_cast<int64_t>(row0) * kpad;
  const int8_t* Bt = d.RB + static_cast<int64_t>(g) * d.planeB + static_cast<int64_t>(col0) * kpad;
  const int tid = threadIdx.x;
  // byte offsets of the thread's 16-byte chunks inside a tile (same for A and B)
  static_assert(CA == CB, "A and B tiles load the same rows per thread");
  uint32_t voff[CA];
#pragma unroll
  for (int u = 0; u < CA; ++u)
    voff[u] = static_cast<uint32_t>(((tid >> 3) + 32 * u) * kpad + 16 * (tid & 7));

  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wm = wid & 1;
  const int wn = wid >> 1;
  const int l32 = lane & 31;
  const int khalf = lane >> 5;
  const int l16 = lane & 15;
  const int kq = lane >> 4;

  AccV acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; ++f)
#pragma unroll
    for (int q = 0; q < FN; ++q)
#pragma unroll
      for (int r = 0; r < kAccRegs; ++r) acc[f][q][r] = 0;

  int4v ra0[CA], rb0[CB], ra1[CA], rb1[CB];
  // k-block index kb is clamped to the last one: the surplus loads of the
  // pipeline tail re-read it and are never used
  auto load = [&](int4v (&ra)[CA], int4v (&rb)[CB], int kb) {
    const int64_t k0 = kBegin + static_cast<int64_t>(min(kb, nk - 1)) * BK;
    const int8_t* Ak = At + k0;
    const int8_t* Bk = Bt + k0;
#pragma unroll
    for (int u = 0; u < CA; ++u) ra[u] = *reinterpret_cast<const int4v*>(Ak + voff[u]);
#pragma unroll
    for (int u = 0; u < CB; ++u) rb[u] = *reinterpret_cast<const int4v*>(Bk + voff[u]);
  };
  auto store = [&](const int4v (&ra)[CA], const int4v (&rb)[CB], int buf) {
#pragma unroll
    for (int u = 0; u < CA; ++u)
      *reinterpret_cast<int4v*>(&lA[buf][(tid >> 3) + 32 * u][16 * (tid & 7)]) = ra[u];
#pragma unroll
    for (int u = 0; u < CB; ++u)
      *reinterpret_cast<int4v*>(&lB[buf][(tid >> 3) + 32 * u][16 * (tid & 7)]) = rb[u];
  };
  // One k-block step over LDS buffer buf: 4 sub-steps of 32 bytes, FM x FN
  // MFMAs each. A single wave per SIMD issues at most one instruction every few
  // cycles, so memory instructions are spread one per MFMA (never in bursts
  // that would leave the matrix pipe idle while they issue):
  //   * LDS reads of the next sub-step's fragments follow the first FM + FN
  //     MFMAs, B first and then A in the order the MFMAs consume them;
  //   * global loads of k-block kb into register set (la, lb), each with its
  //     address computation, follow every other MFMA of the first half;
  //   * the LDS writes of register set (sa, sb) into buffer sbuf follow the last
  //     MFMAs of sub-step 2 and the first of sub-step 3, leaving them half a
  //     sub-step to complete before the barrier.
  // The first sub-step's fragments are read right after the barrier.
  auto step = [&](int buf, int4v (&la)[CA], int4v (&lb)[CB], int kb, const int4v (&sa)[CA],
                  const int4v (&sb)[CB], int sbuf) {
   if constexpr (!M16) {
    constexpr int kSub = BK / 32;          // sub-steps of 32 bytes
    constexpr int kMfma = FM * FN;         // MFMAs per sub-step
    constexpr int kReads = FM + FN;        // LDS fragment reads per sub-step
    constexpr int kLoads = (CA + CB) / kSub;  // global loads per sub-step
    constexpr int kHalfWrites = (CA + CB) / 2;
    static_assert(FM == FN && CA == CB && kSub == 4 && kReads <= kMfma && 2 * kLoads <= kMfma &&
                      kHalfWrites <= kMfma,
                  "tile schedule");
    const int64_t k0 = kBegin + static_cast<int64_t>(min(kb, nk - 1)) * BK;
    const int8_t* Ak = At + k0;
    const int8_t* Bk = Bt + k0;
    int4v a[2][FM], b[2][FN];
    auto read_frags = [&](int set, int ks) {
#pragma unroll
      for (int q = 0; q < FN; ++q)
        b[set][q] = *reinterpret_cast<const int4v*>(
            &lB[buf][wn * WN + 32 * q + l32][ks * 32 + 16 * khalf]);
#pragma unroll
      for (int f = 0; f < FM; ++f)
        a[set][f] = *reinterpret_cast<const int4v*>(
            &lA[buf][wm * WM + 32 * f + l32][ks * 32 + 16 * khalf]);
    };
    read_frags(0, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, kReads, 0);  // LDS reads
#pragma unroll
    for (int ks = 0; ks < kSub; ++ks) {
      const int cur = ks & 1;
#pragma unroll
      for (int h = 0; h < FM; ++h)
#pragma unroll
        for (int q = 0; q < FN; ++q)
          acc[h][q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[cur][h], b[cur][q], acc[h][q], 0, 0, 0);
      if (ks + 1 < kSub) read_frags(cur ^ 1, ks + 1);
#pragma unroll
      for (int j = 0; j < kLoads; ++j) {
        const int gi = kLoads * ks + j;
        if (gi < CA)
          la[gi] = *reinterpret_cast<const int4v*>(Ak + voff[gi]);
        else
          lb[gi - CA] = *reinterpret_cast<const int4v*>(Bk + voff[gi - CA]);
      }
      if (ks >= kSub - 2) {
#pragma unroll
        for (int j = 0; j < kHalfWrites; ++j) {
          const int u = kHalfWrites * (ks - (kSub - 2)) + j;
          if (u < CA)
            *reinterpret_cast<int4v*>(&lA[sbuf][(tid >> 3) + 32 * u][16 * (tid & 7)]) = sa[u];
          else
            *reinterpret_cast<int4v*>(&lB[sbuf][(tid >> 3) + 32 * (u - CA)][16 * (tid & 7)]) =
                sb[u - CA];
        }
      }
      // issue order of this sub-step (masks: 0x008 MFMA, 0x100 LDS read,
      // 0x002 VALU, 0x020 global read, 0x200 LDS write)
#pragma unroll
      for (int i = 0; i < kMfma; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        // with room for it, the LDS reads go two per MFMA right at the start (the
        // earlier they are issued the less the next sub-step waits), the global
        // loads after every other MFMA that follows
        constexpr bool kEarly = kReads == 8 && kMfma == 16;
        const bool early = kEarly && i < kReads / 2;
        const bool loadSlot = kEarly ? (i >= kReads / 2 && i % 2 == 0 && (i - kReads / 2) / 2 < kLoads)
                                     : (i % 2 == 0 && i / 2 < kLoads);
        if (ks + 1 < kSub && early) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
        if (ks + 1 < kSub && !kEarly && i < kReads) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        if (loadSlot) {
          __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
        }
        if ((ks == kSub - 2 && i >= kMfma - kHalfWrites) || (ks == kSub - 1 && i < kHalfWrites))
          __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
      }
    }
   } else {
    // 16 x 16 x 64 MFMAs: two sub-steps of 64 bytes. Sub-step 0 runs by
    // fragment rows of A, sub-step 1 by fragment columns of B, so one register
    // set per operand suffices: A fragment h takes its sub-step-1 bytes once
    // row h is issued, B fragment q once the last row has used it, and
    // sub-step 1 reads nothing from LDS. The MFMAs are volatile inline
    // assembly, so the memory instructions issue where the source puts them:
    //   * before the first MFMA: A fragment 0, then every B fragment (each
    //     MFMA of row 0 waits only for its own B fragment);
    //   * sub-step 0: A fragment q + 1 after MFMA (0, q); after MFMA (h, 0) the
    //     sub-step-1 read of A fragment h - 1; in the last row the sub-step-1
    //     read of B fragment q after MFMA (h, q), and the last A read after the
    //     last MFMA; two global loads of k-block kb into (la, lb) per row;
    //   * sub-step 1: the LDS writes of register set (sa, sb) into buffer sbuf,
    //     one per MFMA, ending a quarter of the sub-step (large tiles) or four
    //     MFMAs before the barrier.
    constexpr int F = FM;
    constexpr int kMfma = F * F;
    constexpr int kLoads = CA + CB;
    constexpr int kWrites = CA + CB;
    constexpr int kWriteStart = F >= 8 ? kMfma - kMfma / 4 - kWrites : kMfma - 4 - kWrites;
    static_assert(FM == FN && CA == CB && BK == 128 && kLoads == 2 * F && F % 4 == 0 &&
                      kWriteStart >= 0,
                  "tile schedule");
    const int64_t k0 = kBegin + static_cast<int64_t>(min(kb, nk - 1)) * BK;
    const int8_t* Ak = At + k0;
    const int8_t* Bk = Bt + k0;
    auto fragA = [&](int f, int ks) {
      return *reinterpret_cast<const int4v*>(&lA[buf][wm * WM + 16 * f + l16][ks * 64 + 16 * kq]);
    };
    auto fragB = [&](int q, int ks) {
      return *reinterpret_cast<const int4v*>(&lB[buf][wn * WN + 16 * q + l16][ks * 64 + 16 * kq]);
    };
    // global load index after MFMA (h, q) of sub-step 0, or -1: two per row
    auto loadIndex = [](int h, int q) {
      return (q % (F / 2) == F / 2 - 1) ? 2 * h + q / (F / 2) : -1;
    };
    int4v a[F], b[F];
    a[0] = fragA(0, 0);
#pragma unroll
    for (int q = 0; q < F; ++q) b[q] = fragB(q, 0);
#pragma unroll
    for (int h = 0; h < F; ++h) {
#pragma unroll
      for (int q = 0; q < F; ++q) {
        i8_mfma16(acc[h][q], a[h], b[q]);
        if (h == 0 && q + 1 < F) a[q + 1] = fragA(q + 1, 0);
        if (q == 0 && h >= 1) a[h - 1] = fragA(h - 1, 1);
        if (h == F - 1) b[q] = fragB(q, 1);
        const int gi = loadIndex(h, q);
        if (gi >= 0) {
          if (gi < CA)
            la[gi] = *reinterpret_cast<const int4v*>(Ak + voff[gi]);
          else
            lb[gi - CA] = *reinterpret_cast<const int4v*>(Bk + voff[gi - CA]);
        }
      }
    }
    a[F - 1] = fragA(F - 1, 1);
#pragma unroll
    for (int q = 0; q < F; ++q) {
#pragma unroll
      for (int h = 0; h < F; ++h) {
        i8_mfma16(acc[h][q], a[h], b[q]);
        const int u = q * F + h - kWriteStart;
        if (u >= 0 && u < kWrites) {
          if (u < CA)
            *reinterpret_cast<int4v*>(&lA[sbuf][(tid >> 3) + 32 * u][16 * (tid & 7)]) = sa[u];
          else
            *reinterpret_cast<int4v*>(&lB[sbuf][(tid >> 3) + 32 * (u - CA)][16 * (tid & 7)]) =
                sb[u - CA];
        }
      }
    }
   }
  };

  // nk is even (splits are whole multiples of 2 * BK) and positive for every
  // launched block; the test only guards the address arithmetic
  if constexpr (M16) {
    // Every launched block has a positive, even number of k-blocks (splits are
    // whole multiples of 2 * BK), so the loop runs and the accumulators reach
    // the epilogue without a join with an untouched path (a join would make
    // the compiler move them before the wait states); the test only guards the
    // address arithmetic.
    if (nk <= 0) return;
    load(ra0, rb0, 0);
    store(ra0, rb0, 0);
    if constexpr (FM >= 8) {
      // Large tiles: the LDS serves ~94 B/clk per CU of the ~128 it can, so its
      // reads and writes are spread evenly over the MFMAs. Fragment set X
      // holds a k-block's sub-step-0 fragments, set Y its sub-step-1 ones;
      // one register set (ra0, rb0) stages the k-block after next. One step,
      // k-block kb in LDS buffer buf (X already read):
      //   * sub-step 0 (MFMAs on X): Y read from buf, one fragment after every
      //     4th MFMA; the staged k-block kb + 1 written to buf ^ 1 (first
      //     half of it), each register reloaded with k-block kb + 2 four
      //     MFMAs after its write;
      //   * sub-step 1 (MFMAs on Y): the second half written and reloaded over
      //     the first 16 MFMAs, then the one barrier of the k-block, then X
      //     read from buf ^ 1 (k-block kb + 1) after every third MFMA.
      // Every read of buf precedes the barrier of its k-block and every write
      // of buf follows it, so one barrier per k-block orders both buffers.
      int4v xa[FM], xb[FN], ya[FM], yb[FN];
      // fragment j of a set in the order the row-major MFMAs consume them:
      // A 0, B 0..F-1, A 1..F-1
      auto readSet = [&](int4v (&sa_)[FM], int4v (&sb_)[FN], int j, int lbuf, int ks) {
        const int8_t* baseA = &lA[lbuf][wm * WM + l16][ks * 64 + 16 * kq];
        const int8_t* baseB = &lB[lbuf][wn * WN + l16][ks * 64 + 16 * kq];
        if (j == 0)
          sa_[0] = *reinterpret_cast<const int4v*>(baseA);
        else if (j <= FN)
          sb_[j - 1] = *reinterpret_cast<const int4v*>(baseB + (j - 1) * 16 * kStride);
        else
          sa_[j - FN] = *reinterpret_cast<const int4v*>(baseA + (j - FN) * 16 * kStride);
      };
      auto writeStaged = [&](int u, int sbuf) {
        if (u < CA)
          *reinterpret_cast<int4v*>(&lA[sbuf][(tid >> 3) + 32 * u][16 * (tid & 7)]) = ra0[u];
        else
          *reinterpret_cast<int4v*>(&lB[sbuf][(tid >> 3) + 32 * (u - CA)][16 * (tid & 7)]) =
              rb0[u - CA];
      };
      auto loadStaged = [&](int u, int kb) {
        const int64_t k0 = kBegin + static_cast<int64_t>(min(kb, nk - 1)) * BK;
        if (u < CA)
          ra0[u] = *reinterpret_cast<const int4v*>(At + k0 + voff[u]);
        else
          rb0[u - CA] = *reinterpret_cast<const int4v*>(Bt + k0 + voff[u - CA]);
      };
      constexpr int kSet = FM + FN;  // fragments per set
      static_assert(FM == FN && CA == CB && CA + CB == 16 && kSet == 16 && FM * FN == 64,
                    "large-tile schedule");
      auto stepF = [&](int buf, int kb) {
#pragma unroll
        for (int i = 0; i < 64; ++i) {
          i8_mfma16(acc[i / FN][i % FN], xa[i / FN], xb[i % FN]);
          if (i % 4 == 0) readSet(ya, yb, i / 4, buf, 1);
          if (i % 8 == 2) writeStaged(i / 8, buf ^ 1);
          if (i % 8 == 6) loadStaged(i / 8, kb + 2);
        }
#pragma unroll
        for (int i = 0; i < 64; ++i) {
          i8_mfma16(acc[i / FN][i % FN], ya[i / FN], yb[i % FN]);
          if (i < 16 && i % 2 == 0) writeStaged(8 + i / 2, buf ^ 1);
          if (i < 16 && i % 2 == 1) loadStaged(8 + i / 2, kb + 2);
          if (i == 15) __syncthreads();
          if (i >= 16 && (i - 16) % 3 == 0) readSet(xa, xb, (i - 16) / 3, buf ^ 1, 0);
        }
      };
      load(ra0, rb0, 1);
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kSet; ++j) readSet(xa, xb, j, 0, 0);
      for (int it = 0; it < nk; it += 2) {
        stepF(0, it);
        stepF(1, it + 1);
      }
    } else {
      load(ra1, rb1, 1);
      __syncthreads();
      for (int it = 0; it < nk; it += 2) {
        step(0, ra0, rb0, it + 2, ra1, rb1, 1);
        __syncthreads();
        step(1, ra1, rb1, it + 3, ra0, rb0, 0);
        __syncthreads();
      }
    }
    // the last MFMAs (inline assembly, 8 passes) finish before any accumulator
    // is read or moved: the wait states come first among the volatile
    // statements, and every later read depends on one of the empty ones
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7\n\ts_nop 7" ::: "memory");
#pragma unroll
    for (int f = 0; f < FM; ++f)
#pragma unroll
      for (int q = 0; q < FN; ++q) asm volatile("" : "+a"(acc[f][q]));
  } else if (nk > 0) {
    load(ra0, rb0, 0);
    store(ra0, rb0, 0);
    load(ra1, rb1, 1);
    __syncthreads();
    for (int it = 0; it < nk; it += 2) {
      // even step: LDS buffer 0 holds k-block it, register set 1 holds it + 1;
      // set 0 receives k-block it + 2
      step(0, ra0, rb0, it + 2, ra1, rb1, 1);
      __syncthreads();
      // odd step: LDS buffer 1 holds k-block it + 1, register set 0 holds it + 2
      step(1, ra1, rb1, it + 3, ra0, rb0, 0);
      __syncthreads();
    }
  }

  // D layout of v_mfma_i32_32x32x32_i8: lane holds column lane % 32 and rows
  // 8 * (r / 4) + 4 * (lane / 32) + r % 4, r = 0..15; of v_mfma_i32_16x16x64_i8:
  // column lane % 16 and rows 4 * (lane / 16) + r, r = 0..3. Four consecutive
  // rows per group of four registers, stored as one word (biased residues) or
  // one 16-byte vector (int32 sums). Output columns hold ldo >= every tile row.
  const int tmod = g / d.products;
  const double dm = d.ms.m[tmod], dinv = d.ms.inv[tmod];
  const double tie = d.ms.kc[tmod].tie;  // m = 256: ties go up, r in [-128, 127]
  constexpr int kFrag = M16 ? 16 : 32;
#pragma unroll
  for (int f = 0; f < FM; ++f) {
#pragma unroll
    for (int q = 0; q < FN; ++q) {
      const int col = col0 + wn * WN + kFrag * q + (M16 ? l16 : l32);
      if (col >= d.n) continue;
#pragma unroll
      for (int rb4 = 0; rb4 < kAccRegs / 4; ++rb4) {
        const int row = M16 ? row0 + wm * WM + 16 * f + 4 * kq
                            : row0 + wm * WM + 32 * f + 8 * rb4 + 4 * khalf;
        if (OUT8) {
          uint32_t w = 0;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // symmetric residue of the INT32 sum: |v| < 2^31, q = round(v / m) exact
            const double v = acc[f][q][4 * rb4 + r];
            const double rr = fma(-rint(fma(v, dinv, tie)), dm, v);
            w |= (static_cast<uint32_t>(static_cast<int>(rr) + 128) & 0xffu) << (8 * r);
          }
          *reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(d.out) +
                                       static_cast<int64_t>(g) * d.outPlane +
                                       static_cast<int64_t>(col) * d.ldo + row) = w;
        } else {
          int4v v;
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = acc[f][q][4 * rb4 + r];
          *reinterpret_cast<int4v*>(static_cast<int*>(d.out) +
                                    static_cast<int64_t>(s) * d.outSplitStride +
                                    static_cast<int64_t>(g) * d.outPlane +
                                    static_cast<int64_t>(col) * d.ldo + row) = v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Chinese remainder reconstruction and FP64 epilogue

// Per output element and modulus: the symmetric residues of Re and Im of the
// product (complex: Re = T1 - T2, Im = T3 - T1 - T2 modulo m), |r| <= m / 2.
template <bool CPLX, bool IN8>
__device__ __forceinline__ void product_residues(const CrtDesc& d, int t, int64_t off,
                                                 double (&re)[4], double (&im)[4]) {
  const double m = d.ms.m[t], inv = d.ms.inv[t];
  double T[3][4];
#pragma unroll
  for (int p = 0; p < (CPLX ? 3 : 1); ++p) {
    const int64_t g = static_cast<int64_t>(t) * d.products + p;
    if (IN8) {
      // biased residues r + 128 of four consecutive rows
      const uint32_t w = *reinterpret_cast<const uint32_t*>(
          static_cast<const uint8_t*>(d.W) + g * d.plane + off);
#pragma unroll
      for (int u = 0; u < 4; ++u) T[p][u] = static_cast<double>(static_cast<int>((w >> (8 * u)) & 0xffu) - 128);
    } else {
      const int* w = static_cast<const int*>(d.W);
#pragma unroll
      for (int u = 0; u < 4; ++u) T[p][u] = 0.0;
      for (int s = 0; s < d.splits; ++s) {
        const int4v v = *reinterpret_cast<const int4v*>(
            w + (static_cast<int64_t>(s) * d.G + g) * d.plane + off);
#pragma unroll
        for (int u = 0; u < 4; ++u) T[p][u] += static_cast<double>(v[u]);  // exact: < 2^38
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (CPLX) {
      re[u] = crt::sym_residue(T[0][u] - T[1][u], m, inv);
      im[u] = crt::sym_residue(T[2][u] - T[0][u] - T[1][u], m, inv);
    } else {
      re[u] = crt::sym_residue(T[0][u], m, inv);
    }
  }
}

// The same from preloaded words of biased residues r + 128 (four consecutive
// rows per word, one word per complex sub-product).
template <bool CPLX>
__device__ __forceinline__ void residues_from_words(const uint32_t* w, double m, double inv,
                                                    double (&re)[4], double (&im)[4]) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    double T[3];
#pragma unroll
    for (int p = 0; p < (CPLX ? 3 : 1); ++p)
      T[p] = static_cast<double>(static_cast<int>((w[p] >> (8 * u)) & 0xffu) - 128);
    if (CPLX) {
      re[u] = crt::sym_residue(T[0] - T[1], m, inv);
      im[u] = crt::sym_residue(T[2] - T[0] - T[1], m, inv);
    } else {
      re[u] = crt::sym_residue(T[0], m, inv);
    }
  }
}

template <bool CPLX>
__device__ __forceinline__ void update_c(const CrtDesc& d, int64_t i, int64_t j, double re,
                                         double im) {
  if (CPLX) {
    double* c = d.C + (j * d.ldc + i) * 2;
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
    double* c = d.C + j * d.ldc + i;
    double out = d.alphaRe * re;
    if (!d.betaZero) out += d.betaRe * c[0];
    c[0] = out;
  }
}

// One thread per four consecutive rows of one column of C: limb sums of the
// CRT over all moduli (crt_arith.hpp), canonical representative, scaling by
// 2^-(ea_i + eb_j) and the alpha / beta update.
template <bool CPLX, bool IN8, int L>
__global__ void __launch_bounds__(256) crt_kernel(CrtDesc d) {
  const int64_t quads = (d.m + 3) / 4;
  const int64_t qid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (qid >= quads * d.n) return;
  const int64_t j = qid / quads;
  const int64_t i0 = 4 * (qid - j * quads);
  if (d.fillMode != 0) {
    const int64_t r0 = (i0 / d.tileRows) * static_cast<int64_t>(d.tileRows);
    const int64_t c0 = (j / d.tileCols) * static_cast<int64_t>(d.tileCols);
    if (!tile_active(d.fillMode, d.fillRowOff, d.fillColOff, r0, c0,
                     min<int64_t>(d.tileRows, d.m - r0), min<int64_t>(d.tileCols, d.n - c0)))
      return;
  }
  const CrtConstants& cc = *d.cc;
  const int64_t off = j * d.ldo + i0;
  double Sre[4][L], Sim[4][L];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int l = 0; l < L; ++l) Sre[u][l] = Sim[u][l] = 0.0;
  auto accumulate = [&](int t, const double (&re)[4], const double (&im)[4]) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const double w = cc.Wl[t][l];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        Sre[u][l] = fma(re[u], w, Sre[u][l]);  // exact (crt_arith.hpp)
        if (CPLX) Sim[u][l] = fma(im[u], w, Sim[u][l]);
      }
    }
  };
  if constexpr (IN8) {
    // every residue word is loaded before any is used, so the loads of all
    // moduli are in flight together
    constexpr int P = CPLX ? 3 : 1;
    uint32_t words[kMaxModuli][P];
#pragma unroll
    for (int t = 0; t < kMaxModuli; ++t) {
      if (t < d.ms.count) {
#pragma unroll
        for (int p = 0; p < P; ++p)
          words[t][p] = *reinterpret_cast<const uint32_t*>(
              static_cast<const uint8_t*>(d.W) +
              (static_cast<int64_t>(t) * d.products + p) * d.plane + off);
      }
    }
#pragma unroll
    for (int t = 0; t < kMaxModuli; ++t) {
      if (t < d.ms.count) {
        double re[4], im[4];
        residues_from_words<CPLX>(words[t], d.ms.m[t], d.ms.inv[t], re, im);
        accumulate(t, re, im);
      }
    }
  } else {
    for (int t = 0; t < d.ms.count; ++t) {
      double re[4], im[4];
      product_residues<CPLX, false>(d, t, off, re, im);
      accumulate(t, re, im);
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t i = i0 + u;
    if (i >= d.m) break;
    const int sc = -(d.expA[i] + d.expB[j]);
    const double re = ldexp(crt::finish<L>(Sre[u], cc), sc);
    const double im = CPLX ? ldexp(crt::finish<L>(Sim[u], cc), sc) : 0.0;
    update_c<CPLX>(d, i, j, re, im);
  }
}

}  // namespace

hipError_t launch_split_stats(bool complex, int layout, const double* X, int64_t ld, int rows,
                              int64_t k, double* partial, int splits, double* outMax,
                              double* outSum, hipStream_t stream) {
  if (rows == 0) return hipSuccess;
  double* pmax = partial;
  double* psum = partial + static_cast<int64_t>(splits) * rows;
  if (layout == KMajor) {
    const int64_t Kp = complex ? 2 * k : k;
    const int64_t per = (Kp + splits - 1) / splits;
    const dim3 grid(rows, splits);
    if (complex)
      hipLaunchKernelGGL(stats_kmajor_kernel<true>, grid, dim3(256), 0, stream, X, ld, rows, Kp,
                         per, pmax, psum);
    else
      hipLaunchKernelGGL(stats_kmajor_kernel<false>, grid, dim3(256), 0, stream, X, ld, rows, Kp,
                         per, pmax, psum);
  } else {
    const int64_t per = (k + splits - 1) / splits;
    const dim3 grid((rows + 63) / 64, splits);
    if (complex)
      hipLaunchKernelGGL(stats_mnmajor_kernel<true>, grid, dim3(256), 0, stream, X, ld, rows, k,
                         per, pmax, psum);
    else
      hipLaunchKernelGGL(stats_mnmajor_kernel<false>, grid, dim3(256), 0, stream, X, ld, rows, k,
                         per, pmax, psum);
  }
  hipLaunchKernelGGL(stats_finalize_kernel, dim3((rows + 255) / 256), dim3(256), 0, stream, pmax,
                     psum, rows, splits, outMax, outSum);
  return hipGetLastError();
}

hipError_t launch_residues(bool complex, int layout, const ResidueDesc& d, hipStream_t stream) {
  if (d.rows == 0) return hipSuccess;
  if (layout == KMajor) {
    const dim3 grid(static_cast<unsigned>((d.kpad + kResElems - 1) / kResElems), d.rows);
    if (complex)
      hipLaunchKernelGGL(residue_kmajor_kernel<true>, grid, dim3(256), 0, stream, d);
    else
      hipLaunchKernelGGL(residue_kmajor_kernel<false>, grid, dim3(256), 0, stream, d);
  } else {
    const dim3 grid(static_cast<unsigned>(d.kpad / kResMnK),
                    (d.rows + kResMnRows - 1) / kResMnRows);
    if (complex)
      hipLaunchKernelGGL(residue_mnmajor_kernel<true>, grid, dim3(256), 0, stream, d);
    else
      hipLaunchKernelGGL(residue_mnmajor_kernel<false>, grid, dim3(256), 0, stream, d);
  }
  return hipGetLastError();
}

// SPLA_AMD_I8_MFMA=32 selects the 32 x 32 x 32 form of the INT8 GEMM (bit-identical
// results, ~10 % slower on the 2048 x 2048 x 2e5 complex product); the default is
// the 16 x 16 x 64 form. Read at every launch.
static bool i8_mfma16() {
  const char* v = std::getenv("SPLA_AMD_I8_MFMA");
  return !(v && std::atoi(v) == 32);
}

hipError_t launch_i8_gemm(const I8GemmDesc& d, int cfg, hipStream_t stream) {
  const unsigned blocks = static_cast<unsigned>(d.tilesM) * d.tilesN * d.G * d.splits;
  const bool out8 = d.splits == 1;
  const bool m16 = i8_mfma16();
  if (cfg == I8TileM) {
    if (m16) {
      if (out8)
        hipLaunchKernelGGL((i8_gemm_kernel<true, 64, 64, true>), dim3(blocks), dim3(256), 0, stream, d);
      else
        hipLaunchKernelGGL((i8_gemm_kernel<false, 64, 64, true>), dim3(blocks), dim3(256), 0, stream, d);
    } else {
      if (out8)
        hipLaunchKernelGGL((i8_gemm_kernel<true, 64, 64, false>), dim3(blocks), dim3(256), 0, stream, d);
      else
        hipLaunchKernelGGL((i8_gemm_kernel<false, 64, 64, false>), dim3(blocks), dim3(256), 0, stream, d);
    }
  } else {
    if (m16) {
      if (out8)
        hipLaunchKernelGGL((i8_gemm_kernel<true, 128, 128, true>), dim3(blocks), dim3(256), 0, stream, d);
      else
        hipLaunchKernelGGL((i8_gemm_kernel<false, 128, 128, true>), dim3(blocks), dim3(256), 0, stream, d);
    } else {
      if (out8)
        hipLaunchKernelGGL((i8_gemm_kernel<true, 128, 128, false>), dim3(blocks), dim3(256), 0, stream, d);
      else
        hipLaunchKernelGGL((i8_gemm_kernel<false, 128, 128, false>), dim3(blocks), dim3(256), 0, stream, d);
    }
  }
  return hipGetLastError();
}

template <bool CPLX, bool IN8>
void launch_crt_limbs(const CrtDesc& d, int limbs, dim3 grid, hipStream_t stream) {
  switch (limbs) {
    case 1: hipLaunchKernelGGL((crt_kernel<CPLX, IN8, 1>), grid, dim3(256), 0, stream, d); break;
    case 2: hipLaunchKernelGGL((crt_kernel<CPLX, IN8, 2>), grid, dim3(256), 0, stream, d); break;
    case 3: hipLaunchKernelGGL((crt_kernel<CPLX, IN8, 3>), grid, dim3(256), 0, stream, d); break;
    default: hipLaunchKernelGGL((crt_kernel<CPLX, IN8, 4>), grid, dim3(256), 0, stream, d); break;
  }
}

hipError_t launch_crt(bool complex, const CrtDesc& d, int limbs, hipStream_t stream) {
  const int64_t total = (static_cast<int64_t>(d.m) + 3) / 4 * d.n;
  if (total == 0) return hipSuccess;
  if (limbs < 1 || limbs > kMaxCrtLimbs || d.ldo % 4 != 0) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>((total + 255) / 256));
  if (complex) {
    if (d.in8)
      launch_crt_limbs<true, true>(d, limbs, grid, stream);
    else
      launch_crt_limbs<true, false>(d, limbs, grid, stream);
  } else {
    if (d.in8)
      launch_crt_limbs<false, true>(d, limbs, grid, stream);
    else
      launch_crt_limbs<false, false>(d, limbs, grid, stream);
  }
  return hipGetLastError();
}

}  // namespace spla_amd
