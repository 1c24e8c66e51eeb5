"""Placement restatement (oracle, test infrastructure only).

Pure-integer functions that decide which blocks of C a ring step of SPLA's
pgemm_ssb / pgemm_ssbtr computes and where every piece of a block lives:

* ``block_is_active``      src/block_generation/block.hpp:58-66
* ``find_optimal_proc_grid`` src/util/block_size_selection.hpp:42-55
* ``block_size_selection`` src/util/block_size_selection.hpp:57-186 (both overloads)
* ``BlockCyclic``          src/block_generation/block_cyclic_generator.cpp:38-112
                           and block_cyclic_generator.hpp:46-50 (sub-generators)
* ``Mirror``               src/block_generation/mirror_generator.hpp:40-110
* ``ssb_ring_blocks``      the grid-wise block order of src/pgemm_ssb/pgemm_ssb_gpu.cpp:101-112
                           (block shape) and :170-190 (order, fill filter)

SPLA's IntType is std::ptrdiff_t; conversions from double truncate toward zero
(``int()`` here), ``std::round`` rounds halves away from zero (``_round``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

FILL_FULL, FILL_UPPER, FILL_LOWER = 0, 1, 2
RING_THRESHOLD = 0.65   # pgemm_ssb_gpu.cpp:109
MIN_BLOCK_SIZE = 375    # pgemm_ssb_gpu.cpp:110


def _round(x: float) -> int:
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def block_is_active(row, col, rows, cols, row_off, col_off, mode) -> bool:
    """block.hpp:58-66."""
    if mode == FILL_UPPER:
        return col + cols + col_off > row + row_off
    if mode == FILL_LOWER:
        return row + rows + row_off > col + col_off
    return True


def find_optimal_proc_grid(comm_size, lower_dev, upper_dev):
    """block_size_selection.hpp:42-55 (deviations are IntType: truncated by the caller)."""
    s = int(math.sqrt(comm_size))
    for rows in range(s, comm_size + 1):
        for cols in range(s, 0, -1):
            if comm_size - lower_dev <= rows * cols <= comm_size + upper_dev:
                return rows, cols
    return comm_size, 1


def _target_shape(m, n, target):
    """The 'use the target block size' branch (block_size_selection.hpp:79-90, 164-175)."""
    r = min(target, m)
    c = min(target, n)
    if r < target:
        c = int(c * (float(target) / float(r)))
    if c < target:
        r = int(r * (float(target) / float(c)))
    return min(r, m), min(c, n)


def block_size_selection_full(disjoint, deviation, comm_size, m, n, target, min_block):
    """block_size_selection.hpp:57-131."""
    if m * n <= min_block * min_block:
        return m, n
    if not disjoint:
        return min(target, m), min(target, n)
    gr, gc = find_optimal_proc_grid(comm_size, int(deviation * comm_size), 0)
    if m > n and gr < gc:
        gr, gc = gc, gr
    rib = (m + gr - 1) // gr
    cib = (n + gc - 1) // gc
    if rib * cib < min_block * min_block:
        return _target_shape(m, n, target)
    factor = float(rib * cib) / float(target * target)
    if factor >= 1.5:
        rf = int(math.sqrt(factor))
        cf = int(math.ceil(factor / rf))
        if m > n and rf < cf:
            rf, cf = cf, rf
        gr *= rf
        gc *= cf
        rib = (m + gr - 1) // gr
        cib = (n + gc - 1) // gc
    excess = (gr * gc) % comm_size
    if excess > 0:
        missing = comm_size - excess
        if gr > gc:
            if gr <= missing:
                gc += 1
            elif gc <= missing:
                gr += 1
        else:
            if gc <= missing:
                gr += 1
            elif gr <= missing:
                gc += 1
        rib = (m + gr - 1) // gr
        cib = (n + gc - 1) // gc
    return rib, cib


def block_size_selection(mode, disjoint, deviation, comm_size, m, n, row_off, col_off, target,
                         min_block):
    """block_size_selection.hpp:133-186."""
    rib, cib = block_size_selection_full(disjoint, deviation, comm_size, m, n, target, min_block)
    if mode != FILL_FULL and disjoint:
        nbr = (m + rib - 1) // rib
        nbc = (n + cib - 1) // cib
        active = 0
        for r in range(0, m, rib):
            for c in range(0, n, cib):
                if block_is_active(r, c, min(rib, m - r), min(cib, n - c), row_off, col_off, mode):
                    active += 1
        # no active block: the shape is kept (the reference would divide by zero here)
        factor = float(nbr * nbc) / float(active) if active else 1.0
        if factor > 1.0:
            factor = math.sqrt(factor)
            nbr = int(math.ceil(factor * nbr))
            nbc = _round(factor * nbc)
            rib = (m + nbr - 1) // nbr
            cib = (n + nbc - 1) // nbc
            if rib * cib < min_block * min_block:
                rib, cib = _target_shape(m, n, target)
    return rib, cib


@dataclass(frozen=True)
class Piece:
    """BlockInfo (block.hpp:40-48)."""
    global_row: int
    global_col: int
    sub_row: int
    sub_col: int
    local_row: int
    local_col: int
    rows: int
    cols: int
    rank: int


class BlockCyclic:
    """BlockCyclicGenerator (block_cyclic_generator.cpp:38-112)."""

    disjoint = True

    def __init__(self, rib, cib, grid_rows, grid_cols, rows, cols, row_off, col_off):
        self.rib, self.cib = rib, cib
        self.gr, self.gc = grid_rows, grid_cols
        self.rows, self.cols = rows, cols
        self.row_off, self.col_off = row_off, col_off
        first_r = (row_off // rib) * rib
        first_c = (col_off // cib) * cib
        self.nbr = (row_off - first_r + rows + rib - 1) // rib
        self.nbc = (col_off - first_c + cols + cib - 1) // cib

    def sub(self, row, col, rows, cols):
        """block_cyclic_generator.hpp:46-50."""
        return BlockCyclic(self.rib, self.cib, self.gr, self.gc, rows, cols, row + self.row_off,
                           col + self.col_off)

    def num_blocks(self):
        return self.nbr * self.nbc

    def info(self, idx) -> Piece:
        br = idx % self.nbr
        bc = idx // self.nbr
        gbr = br + self.row_off // self.rib
        gbc = bc + self.col_off // self.cib
        first_r = (self.row_off // self.rib) * self.rib
        first_c = (self.col_off // self.cib) * self.cib
        grow = max(self.row_off, first_r + br * self.rib)
        gcol = max(self.col_off, first_c + bc * self.cib)
        rows = min((gbr + 1) * self.rib, self.row_off + self.rows) - grow
        cols = min((gbc + 1) * self.cib, self.col_off + self.cols) - gcol
        rank = gbr % self.gr + (gbc % self.gc) * self.gr
        lrow = (gbr // self.gr) * self.rib + grow % self.rib
        lcol = (gbc // self.gc) * self.cib + gcol % self.cib
        return Piece(grow, gcol, grow - self.row_off, gcol - self.col_off, lrow, lcol, rows, cols,
                     rank)


class Mirror:
    """MirrorGenerator (mirror_generator.hpp:40-110): every rank holds all of C."""

    disjoint = False

    def __init__(self, rib, cib, rows, cols, row_off, col_off):
        self.rib, self.cib = rib, cib
        self.rows, self.cols = rows, cols
        self.row_off, self.col_off = row_off, col_off
        self.nbr = (rows + rib - 1) // rib
        self.nbc = (cols + cib - 1) // cib

    def sub(self, row, col, rows, cols):
        return Mirror(self.rib, self.cib, rows, cols, row + self.row_off, col + self.col_off)

    def num_blocks(self):
        return self.nbr * self.nbc

    def info(self, idx) -> Piece:
        br = idx % self.nbr
        bc = idx // self.nbr
        grow = br * self.rib + self.row_off
        gcol = bc * self.cib + self.col_off
        rows = min(self.rows - br * self.rib, self.rib)
        cols = min(self.cols - bc * self.cib, self.cib)
        return Piece(grow, gcol, grow - self.row_off, gcol - self.col_off, grow, gcol, rows, cols,
                     -1)


def ssb_ring_blocks(mirror: bool, grid_rows, grid_cols, comm_size, m, n, row_off, col_off, mode,
                    tile_size_host):
    """Block shape and ordered active blocks of a pgemm_ssb / ssbtr call
    (pgemm_ssb_gpu.cpp:101-112, 170-190); consecutive groups of comm_size blocks
    form one ring tile. Returns (rows_in_block, cols_in_block, [(row, col, rows, cols)])."""
    if m == 0 or n == 0:  # pgemm_ssb returns before planning (pgemm_ssb_gpu.cpp:247-249)
        return 0, 0, []
    rib, cib = block_size_selection(mode, not mirror, 1.0 - RING_THRESHOLD, comm_size, m, n,
                                    row_off, col_off, tile_size_host, MIN_BLOCK_SIZE)
    blocks = []
    for c0 in range(0, n, grid_cols * cib):
        for r0 in range(0, m, grid_rows * rib):
            for c in range(c0, min(n, c0 + grid_cols * cib), cib):
                for r in range(r0, min(m, r0 + grid_rows * rib), rib):
                    b = (r, c, min(rib, m - r), min(cib, n - c))
                    if block_is_active(*b, row_off, col_off, mode):
                        blocks.append(b)
    return rib, cib, blocks


def block_pieces(mirror: bool, rib, cib, grid_rows, grid_cols, rows, cols, row_off, col_off):
    """Every piece (BlockInfo) of the region rows x cols at (row_off, col_off)."""
    gen = (Mirror(rib, cib, rows, cols, row_off, col_off) if mirror else
           BlockCyclic(rib, cib, grid_rows, grid_cols, rows, cols, row_off, col_off))
    return [gen.info(i) for i in range(gen.num_blocks())]
