// Ring-step planning of pgemm_ssb / pgemm_ssbtr: which blocks of C one ring
// step computes, in which order, and where every piece of a block lives on
// the process grid. Pure integer host code; the block products themselves run
// on the batched engine (batched.hpp).
//
// The rules are SPLA's own, so that a ring step planned here holds exactly the
// blocks SPLA's ring would hold:
//   fill predicate       src/block_generation/block.hpp:58-66
//   block shape          src/util/block_size_selection.hpp:42-186, with the
//                        constants of src/pgemm_ssb/pgemm_ssb_gpu.cpp:108-112
//   block order          src/pgemm_ssb/pgemm_ssb_gpu.cpp:170-190
//   piece placement      src/block_generation/block_cyclic_generator.cpp:38-112,
//                        mirror_generator.hpp:40-110
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace spla_amd {

struct RingBlock {
  int64_t row, col, rows, cols;  // inside the m x n sub-matrix of C
};

// One piece of a region of C on the process grid (SPLA's BlockInfo).
struct PlacedPiece {
  int64_t globalRow, globalCol;  // first element in the global matrix
  int64_t subRow, subCol;        // same, relative to the region's origin
  int64_t localRow, localCol;    // first element in the owner's local matrix
  int64_t rows, cols;
  int64_t rank;  // owner (column-major over the grid); -1: held by every rank
};

bool ring_block_active(int fillMode, int64_t row, int64_t col, int64_t rows, int64_t cols,
                       int64_t rowOff, int64_t colOff);

// Shape of the ring blocks for an m x n product (fill mode and offsets of C).
std::pair<int64_t, int64_t> ring_block_shape(int fillMode, bool disjoint, double deviation,
                                             int64_t commSize, int64_t m, int64_t n,
                                             int64_t rowOff, int64_t colOff, int64_t target,
                                             int64_t minBlock);

// Ordered active blocks of a pgemm_ssb / ssbtr call; consecutive groups of
// commSize blocks form one ring tile.
struct SsbRingPlan {
  int64_t rowsInBlock = 0, colsInBlock = 0;
  std::vector<RingBlock> blocks;
};
SsbRingPlan plan_ssb_ring(bool mirror, int64_t gridRows, int64_t gridCols, int64_t commSize,
                          int64_t m, int64_t n, int64_t rowOff, int64_t colOff, int fillMode,
                          int64_t tileSizeHost);

// Pieces of the region rows x cols at (rowOff, colOff) of C: block-cyclic
// blocks of rowsInBlock x colsInBlock on a gridRows x gridCols grid, or (mirror)
// tiles of that size held by every rank.
std::vector<PlacedPiece> region_pieces(bool mirror, int64_t rowsInBlock, int64_t colsInBlock,
                                       int64_t gridRows, int64_t gridCols, int64_t rows,
                                       int64_t cols, int64_t rowOff, int64_t colOff);

}  // namespace spla_amd
