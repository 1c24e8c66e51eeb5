// Descriptors of the persistent batched kernel (batched.hip): many independent
// block products C_b <- alpha op(A_b) B_b + beta C_b (ring-step block products
// of pgemm_ssb / pgemm_ssbtr) in one launch.
#pragma once

#include <cstdint>

namespace spla_amd {

struct BatchBlock {
  const double* A;
  const double* B;
  double* C;
  int64_t lda, ldb, ldc;  // elements
  int m, n;               // block of C (elements)
  int64_t Kp;             // embedded reduction length (k or 2k)
  int Mp;                 // embedded rows (m or 2m)
  int tilesN;
  int splits;
  int64_t kSplit;         // embedded reduction length per split
  int64_t workOffset;     // doubles: this block's [splits][n][Mp] partial sums
  int counterBase;        // first of the block's tilesM * tilesN arrival counters
};

// One work item: (block, tile row, tile column, split).
struct BatchItem {
  int block, tm, tn, s;
};

struct BatchDesc {
  const BatchBlock* blocks;
  const BatchItem* items;
  int numItems;
  double* work;
  unsigned* counters;  // per split tile; every launch leaves them at 0
  unsigned* queue;     // [0] next item, [1] finished workgroups; left at 0
  double alphaRe, alphaIm, betaRe, betaIm;
  int betaZero;
  int conjA, conjB;
};

}  // namespace spla_amd
