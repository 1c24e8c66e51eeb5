// Host/device shared descriptors of the FP64 and FP32 MFMA GEMM kernels.
//
// The kernels compute C <- alpha * opA(A) * opB(B) + beta * C for column-major
// real or interleaved complex operands of one scalar type (double or float;
// the FP32 kernels accumulate in FP32, as a single-precision BLAS does). A
// complex product is run as a real product of the "real embedding": every
// 16-row MFMA fragment holds the real and the imaginary rows of 8 complex rows
// of C (FP64: rows 0-7 real, 8-15 imaginary; FP32: rows 2c real, 2c + 1
// imaginary, following each instruction's accumulator layout), and the reduction index q = 2*l + c runs over the real (c = 0) and
// imaginary (c = 1) parts of A(l, i) / B(l, j). The complex product is then
// one real GEMM of size (2m) x n x (2k), i.e. exactly 8*m*n*k flops, with no
// operand expanded in memory. Split-K partial sums are stored as interleaved
// (re, im) pairs, i.e. in C's own column-major complex layout.
#pragma once

#include <cstdint>

namespace spla_amd {

// Storage order of an operand as seen by the product.
//   KMajor : the reduction index is contiguous (A^T/A^H read of A, B read of B).
//   MNMajor: the output index is contiguous (A read of A, B^T/B^H read of B).
enum OperandLayout : int { KMajor = 0, MNMajor = 1 };

// Fill predicate of a rectangular tile of C, following SPLA's block_is_active
// (src/block_generation/block.hpp:58-66). Coordinates are in elements of the
// (sub-)matrix C; rowOff / colOff locate it in the global matrix.
enum TileFill : int { FillFull = 0, FillUpper = 1, FillLower = 2 };

struct GemmDesc {
  const void* A;  // first element of A (scalars of the kernel's type)
  const void* B;
  void* C;
  void* work;  // split-K partial sums: [splits][N][Mp], raw (unscaled) sums

  int64_t lda, ldb, ldc;  // leading dimensions in elements (complex or real)

  int m, n;   // output rows / columns in elements
  int64_t k;  // reduction length in elements

  int Mp;         // embedded output rows: m (real) or 2m (complex)
  int64_t Kp;     // embedded reduction length: k (real) or 2k (complex)
  int64_t kSplit; // embedded reduction length handled by one split (multiple of the k-block)
  int splits;
  int tilesM, tilesN;

  double alphaRe, alphaIm, betaRe, betaIm;
  int betaZero;  // beta == 0: C is written without being read

  int conjA, conjB;  // complex conjugation of the operand (complex only)

  int fillMode;  // TileFill
  int64_t fillRowOff, fillColOff;
};

// Tile configurations of the FP64 MFMA kernel. Rows are embedded rows.
enum GemmTileConfig : int {
  TileL = 0,   // 128 x 128 block, 4 waves of 64 x 64, 2 blocks per CU
  TileM = 1,   // 64 x 64 block, 4 waves of 32 x 32
  TileXL = 2,  // 256 x 128 block, 4 waves of 128 x 64, 1 block per CU
  NumTileConfigs = 3
};

struct TileShape {
  int bm, bn, bk;
};

inline TileShape tile_shape(int cfg) {
  switch (cfg) {
    case TileM:
      return {64, 64, 16};
    case TileXL:
      return {256, 128, 16};
    default:
      return {128, 128, 16};
  }
}

}  // namespace spla_amd
