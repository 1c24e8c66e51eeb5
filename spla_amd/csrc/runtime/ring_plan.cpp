#include "runtime/ring_plan.hpp"

#include <algorithm>
#include <cmath>
#include <tuple>

#include "spla/types.h"

namespace spla_amd {
namespace {

constexpr double kRingThreshold = 0.65;  // share of blocks per rank that enables the ring
constexpr int64_t kMinBlock = 375;       // smallest ring block edge (same on every rank)

// Process grid close to commSize (rows >= floor(sqrt), largest cols first)
// with rows * cols in [commSize - lower, commSize + upper].
std::pair<int64_t, int64_t> near_grid(int64_t commSize, int64_t lower, int64_t upper) {
  const int64_t root = static_cast<int64_t>(std::sqrt(static_cast<double>(commSize)));
  for (int64_t r = root; r <= commSize; ++r)
    for (int64_t c = root; c > 0; --c) {
      const int64_t p = r * c;
      if (p >= commSize - lower && p <= commSize + upper) return {r, c};
    }
  return {commSize, 1};
}

// Blocks of the target size, stretched along one edge when the other is short.
std::pair<int64_t, int64_t> target_shape(int64_t m, int64_t n, int64_t target) {
  int64_t r = std::min(target, m), c = std::min(target, n);
  if (r < target) c = static_cast<int64_t>(c * (static_cast<double>(target) / r));
  if (c < target) r = static_cast<int64_t>(r * (static_cast<double>(target) / c));
  return {std::min(r, m), std::min(c, n)};
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

std::pair<int64_t, int64_t> full_shape(bool disjoint, double deviation, int64_t commSize,
                                       int64_t m, int64_t n, int64_t target, int64_t minBlock) {
  if (m * n <= minBlock * minBlock) return {m, n};
  if (!disjoint) return {std::min(target, m), std::min(target, n)};
  auto grid = near_grid(commSize, static_cast<int64_t>(deviation * commSize), 0);
  if (m > n && grid.first < grid.second) std::swap(grid.first, grid.second);
  int64_t rib = ceil_div(m, grid.first), cib = ceil_div(n, grid.second);
  if (rib * cib < minBlock * minBlock) return target_shape(m, n, target);
  const double factor = static_cast<double>(rib * cib) / static_cast<double>(target * target);
  if (factor >= 1.5) {
    // too large: multiply the block count by an integer split of factor
    int64_t rf = static_cast<int64_t>(std::sqrt(factor));
    int64_t cf = static_cast<int64_t>(std::ceil(factor / static_cast<double>(rf)));
    if (m > n && rf < cf) std::swap(rf, cf);
    grid.first *= rf;
    grid.second *= cf;
    rib = ceil_div(m, grid.first);
    cib = ceil_div(n, grid.second);
  }
  // move the block count closer to a multiple of the communicator size
  const int64_t excess = (grid.first * grid.second) % commSize;
  if (excess > 0) {
    const int64_t missing = commSize - excess;
    int64_t& larger = grid.first > grid.second ? grid.first : grid.second;
    int64_t& smaller = grid.first > grid.second ? grid.second : grid.first;
    if (larger <= missing)
      smaller += 1;
    else if (smaller <= missing)
      larger += 1;
    rib = ceil_div(m, grid.first);
    cib = ceil_div(n, grid.second);
  }
  return {rib, cib};
}

}  // namespace

bool ring_block_active(int fillMode, int64_t row, int64_t col, int64_t rows, int64_t cols,
                       int64_t rowOff, int64_t colOff) {
  if (fillMode == SPLA_FILL_MODE_UPPER) return col + cols + colOff > row + rowOff;
  if (fillMode == SPLA_FILL_MODE_LOWER) return row + rows + rowOff > col + colOff;
  return true;
}

std::pair<int64_t, int64_t> ring_block_shape(int fillMode, bool disjoint, double deviation,
                                             int64_t commSize, int64_t m, int64_t n,
                                             int64_t rowOff, int64_t colOff, int64_t target,
                                             int64_t minBlock) {
  auto shape = full_shape(disjoint, deviation, commSize, m, n, target, minBlock);
  if (fillMode == SPLA_FILL_MODE_FULL || !disjoint) return shape;
  // triangular: grow the block count by the share of inactive blocks
  int64_t nbr = ceil_div(m, shape.first), nbc = ceil_div(n, shape.second);
  int64_t active = 0;
  for (int64_t r = 0; r < m; r += shape.first)
    for (int64_t c = 0; c < n; c += shape.second)
      if (ring_block_active(fillMode, r, c, std::min(shape.first, m - r),
                            std::min(shape.second, n - c), rowOff, colOff))
        ++active;
  if (active == 0) return shape;  // nothing to compute: keep the shape
  double factor = static_cast<double>(nbr * nbc) / static_cast<double>(active);
  if (factor > 1.0) {
    factor = std::sqrt(factor);
    nbr = static_cast<int64_t>(std::ceil(factor * nbr));
    nbc = static_cast<int64_t>(std::round(factor * nbc));
    shape = {ceil_div(m, nbr), ceil_div(n, nbc)};
    if (shape.first * shape.second < minBlock * minBlock) shape = target_shape(m, n, target);
  }
  return shape;
}

SsbRingPlan plan_ssb_ring(bool mirror, int64_t gridRows, int64_t gridCols, int64_t commSize,
                          int64_t m, int64_t n, int64_t rowOff, int64_t colOff, int fillMode,
                          int64_t tileSizeHost) {
  SsbRingPlan p;
  if (m == 0 || n == 0) return p;  // pgemm_ssb returns before planning
  std::tie(p.rowsInBlock, p.colsInBlock) =
      ring_block_shape(fillMode, !mirror, 1.0 - kRingThreshold, commSize, m, n, rowOff, colOff,
                       tileSizeHost, kMinBlock);
  const int64_t rib = p.rowsInBlock, cib = p.colsInBlock;
  // grid-wise: one grid of blocks (gridRows x gridCols) after the other,
  // column-major inside the grid
  for (int64_t c0 = 0; c0 < n; c0 += gridCols * cib)
    for (int64_t r0 = 0; r0 < m; r0 += gridRows * rib)
      for (int64_t c = c0; c < std::min(n, c0 + gridCols * cib); c += cib)
        for (int64_t r = r0; r < std::min(m, r0 + gridRows * rib); r += rib) {
          const RingBlock b{r, c, std::min(rib, m - r), std::min(cib, n - c)};
          if (ring_block_active(fillMode, b.row, b.col, b.rows, b.cols, rowOff, colOff))
            p.blocks.push_back(b);
        }
  return p;
}

std::vector<PlacedPiece> region_pieces(bool mirror, int64_t rib, int64_t cib, int64_t gridRows,
                                       int64_t gridCols, int64_t rows, int64_t cols,
                                       int64_t rowOff, int64_t colOff) {
  std::vector<PlacedPiece> out;
  if (mirror) {
    const int64_t nbr = ceil_div(rows, rib), nbc = ceil_div(cols, cib);
    out.reserve(nbr * nbc);
    for (int64_t bc = 0; bc < nbc; ++bc)
      for (int64_t br = 0; br < nbr; ++br) {
        const int64_t gr = br * rib + rowOff, gc = bc * cib + colOff;
        out.push_back({gr, gc, gr - rowOff, gc - colOff, gr, gc,
                       std::min(rows - br * rib, rib), std::min(cols - bc * cib, cib), -1});
      }
    return out;
  }
  // block-cyclic: the region starts inside global block (rowOff / rib, colOff / cib)
  const int64_t firstBr = rowOff / rib, firstBc = colOff / cib;
  const int64_t nbr = ceil_div(rowOff - firstBr * rib + rows, rib);
  const int64_t nbc = ceil_div(colOff - firstBc * cib + cols, cib);
  out.reserve(nbr * nbc);
  for (int64_t bc = 0; bc < nbc; ++bc)
    for (int64_t br = 0; br < nbr; ++br) {
      const int64_t gbr = firstBr + br, gbc = firstBc + bc;
      const int64_t gr = std::max(rowOff, gbr * rib), gc = std::max(colOff, gbc * cib);
      PlacedPiece q;
      q.globalRow = gr;
      q.globalCol = gc;
      q.subRow = gr - rowOff;
      q.subCol = gc - colOff;
      q.rows = std::min((gbr + 1) * rib, rowOff + rows) - gr;
      q.cols = std::min((gbc + 1) * cib, colOff + cols) - gc;
      q.rank = gbr % gridRows + (gbc % gridCols) * gridRows;
      q.localRow = (gbr / gridRows) * rib + gr % rib;
      q.localCol = (gbc / gridCols) * cib + gc % cib;
      out.push_back(q);
    }
  return out;
}

}  // namespace spla_amd
