// Schedule of a batch of block products for the grouped launch (batched.hip):
// the tiles of every block, the reduction splits of each block, the order in
// which the workgroups are dispatched, the split-K workspace layout and a cost
// model of the launch. Pure host code without HIP, so that the schedule and the
// kernels' indexing over it can be checked on the CPU.
//
// Model: every tile costs 2 BM BN max(Kp, BK) flops at the per-workgroup share
// of the FP64 matrix rate (CU rate x efficiency / resident workgroups per CU),
// plus a fixed overhead per workgroup. A block's tiles are split along the
// reduction so that one workgroup costs about a third of a slot's share of the
// batch (at least 20 us), with at least 16 k-blocks per split. Workgroups are
// dispatched in decreasing cost order; the hardware hands them to the slots as
// the slots free up, which the model simulates (longest processing time first).
#pragma once

#include <cstdint>
#include <vector>

#include "kernels/batched_desc.hpp"

namespace spla_amd {

struct BatchShape {
  int64_t m, n, k;
  int fillMode;             // SplaFillMode (ssb blocks; sbs blocks are full)
  int64_t rowOff, colOff;   // position of the block in the global C
};

struct BatchSchedule {
  int cfg = 0;              // GemmTileConfig: TileL or TileM
  double splitScale = 1.0;  // scales the split counts of the model (tuning candidates)
  std::vector<BatchBlock> blocks;  // one per shape; pointers and leading dimensions unset
  std::vector<BatchItem> items;    // GEMM workgroups in dispatch order
  // static schedule of the persistent launch: slot w (one workgroup of the
  // launch) takes items slotItems[slotStart[w] .. slotStart[w + 1]) in order,
  // as the model's longest-processing-time simulation assigned them
  std::vector<int> slotStart, slotItems;
  std::vector<BatchReduceTile> reduce;
  int64_t workDoubles = 0;
  double modelSeconds = 0.0;
};

// Resident workgroups per CU of the batched GEMM kernel of a tile configuration.
int batch_blocks_per_cu(int cfg);
// Workgroups of the persistent launch: the resident slots, at most one per item.
int batch_slots(int cfg, int numCUs, std::size_t items);

// efficiency: the FP64 matrix rate a tile sustains, as a fraction of the peak.
BatchSchedule plan_batch_schedule(const std::vector<BatchShape>& shapes, bool complex,
                                  bool alphaZero, int cfg, double splitScale, int numCUs,
                                  double efficiency);

// Fill predicate of a tile (SPLA's block_is_active, src/block_generation/block.hpp:58-66).
bool batch_tile_active(int fillMode, int64_t rowOff, int64_t colOff, int64_t r, int64_t c,
                       int64_t nr, int64_t nc);

}  // namespace spla_amd
