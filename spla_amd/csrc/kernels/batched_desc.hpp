// Descriptors of the batched block-product kernels (batched.hip): many
// independent block products C_b <- alpha op(A_b) B_b + beta C_b (the block
// products of one pgemm_ssb / pgemm_ssbtr or pgemm_sbs ring step) in one GEMM
// launch over the tiles of every block (grouped: one workgroup per tile and
// split; or persistent: one workgroup per CU slot with a static list of them),
// followed by one fixed-order reduction launch over the tiles whose reduction
// is split.
#pragma once

#include <cstdint>

namespace spla_amd {

struct BatchBlock {
  const double* A;
  const double* B;
  double* C;
  int64_t lda, ldb, ldc;  // elements
  int m, n;               // block of C (elements)
  int64_t Kp;             // embedded reduction length (k or 2k; 0: C <- beta C only)
  int Mp;                 // embedded rows (m or 2m)
  int splits;             // reduction splits of every tile of the block
  int64_t kSplit;         // embedded reduction length per split (whole k-blocks)
  int64_t workOffset;     // doubles: this block's [splits][n][Mp] partial sums
};

// One workgroup of the GEMM launch: (block, tile row, tile column, split).
struct BatchItem {
  int block, tm, tn, s;
};

// One workgroup of the reduction launch: a tile of a block with splits > 1.
struct BatchReduceTile {
  int block, tm, tn;
};

struct BatchDesc {
  const BatchBlock* blocks;
  const BatchItem* items;          // grouped GEMM launch: one workgroup per item
  // persistent GEMM launch: workgroup w works through the items
  // slotItems[slotStart[w] .. slotStart[w + 1]) in order (a static schedule)
  const int* slotStart;
  const int* slotItems;
  const BatchReduceTile* reduce;   // gridDim.x of the reduction launch
  double* work;
  double alphaRe, alphaIm, betaRe, betaIm;
  int betaZero;
  int conjA, conjB;
};

}  // namespace spla_amd
