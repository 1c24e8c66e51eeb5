"""Ring-step planning: block shape, block order and piece placement.

* the library (spla_amd_ssb_ring_blocks, spla_amd_region_pieces) equals the
  restatement oracle/placement.py on thousands of random cases (bit-exact);
* the restatement is pinned by closed forms: ScaLAPACK's INDXG2P / INDXG2L
  for the owner and local index of every element, exact partition of every
  region by its pieces and of the active part of C by the ring blocks, the
  fill requirement of pgemm_ssbtr, and hand-computed block shapes;
* planted mistakes in a copy of the restatement are rejected by those checks.
"""

from __future__ import annotations

import random

import numpy as np
import pytest

import spla_amd as spla
from oracle import placement as P

BC, MIRROR = 0, 1


def random_region(rng):
    rb, cb = rng.randint(1, 40), rng.randint(1, 40)
    gr, gc = rng.randint(1, 4), rng.randint(1, 4)
    rows, cols = rng.randint(0, 150), rng.randint(0, 150)
    ro, co = rng.randint(0, 120), rng.randint(0, 120)
    return rb, cb, gr, gc, rows, cols, ro, co


@pytest.mark.parametrize("seed", range(4))
def test_region_pieces_match_restatement(seed):
    rng = random.Random(seed)
    for _ in range(500):
        mirror = rng.random() < 0.2
        rb, cb, gr, gc, rows, cols, ro, co = random_region(rng)
        got = spla.region_pieces(MIRROR if mirror else BC, rb, cb, gr, gc, rows, cols, ro, co)
        exp = [tuple(vars(p).values()) if hasattr(p, "__dict__") else tuple(p.__dict__.values())
               for p in P.block_pieces(mirror, rb, cb, gr, gc, rows, cols, ro, co)]
        assert got == exp, (mirror, rb, cb, gr, gc, rows, cols, ro, co)


def random_ring_case(rng):
    mirror = rng.random() < 0.2
    gr, gc = rng.randint(1, 4), rng.randint(1, 4)
    comm = gr * gc if not mirror else rng.randint(1, 8)
    m, n = rng.choice([rng.randint(0, 300), rng.randint(300, 6000)]), \
        rng.choice([rng.randint(0, 300), rng.randint(300, 6000)])
    ro, co = rng.randint(0, 3000), rng.randint(0, 3000)
    fill = rng.choice([0, 1, 2])
    tile = rng.choice([64, 150, 375, 500, 1500])
    return mirror, gr, gc, comm, m, n, ro, co, fill, tile


@pytest.mark.parametrize("seed", range(4))
def test_ring_blocks_match_restatement(seed):
    rng = random.Random(100 + seed)
    for _ in range(250):
        mirror, gr, gc, comm, m, n, ro, co, fill, tile = random_ring_case(rng)
        got = spla.ssb_ring_blocks(MIRROR if mirror else BC, gr, gc, comm, m, n, ro, co, fill,
                                   tile)
        rib, cib, blocks = P.ssb_ring_blocks(mirror, gr, gc, comm, m, n, ro, co, fill, tile)
        assert got == (rib, cib, blocks), (mirror, gr, gc, comm, m, n, ro, co, fill, tile)


# ---------------------------------------------------------------------------
# pins of the restatement

def indxg2p(i, nb, nprocs):
    """ScaLAPACK INDXG2P (0-based, source process 0)."""
    return (i // nb) % nprocs


def indxg2l(i, nb, nprocs):
    """ScaLAPACK INDXG2L (0-based)."""
    return (i // (nb * nprocs)) * nb + i % nb


def check_pieces(pieces, mirror, rb, cb, gr, gc, rows, cols, ro, co):
    if rows == 0 or cols == 0:  # an empty region: any piece is empty too
        assert all(p.rows == 0 or p.cols == 0 for p in pieces)
        return
    cover = np.zeros((rows, cols), dtype=np.int64)
    for p in pieces:
        assert p.rows > 0 and p.cols > 0
        assert p.sub_row == p.global_row - ro and p.sub_col == p.global_col - co
        cover[p.sub_row:p.sub_row + p.rows, p.sub_col:p.sub_col + p.cols] += 1
        for di in (0, p.rows - 1):
            for dj in (0, p.cols - 1):
                i, j = p.global_row + di, p.global_col + dj
                if mirror:
                    assert p.rank == -1 and (p.local_row + di, p.local_col + dj) == (i, j)
                else:
                    assert p.rank == indxg2p(i, rb, gr) + indxg2p(j, cb, gc) * gr
                    assert p.local_row + di == indxg2l(i, rb, gr)
                    assert p.local_col + dj == indxg2l(j, cb, gc)
        if not mirror:  # a piece never crosses a block boundary
            assert p.global_row // rb == (p.global_row + p.rows - 1) // rb
            assert p.global_col // cb == (p.global_col + p.cols - 1) // cb
    assert (cover == 1).all()


def test_pieces_pinned_by_scalapack_closed_forms():
    rng = random.Random(7)
    for _ in range(400):
        mirror = rng.random() < 0.2
        rb, cb, gr, gc, rows, cols, ro, co = random_region(rng)
        check_pieces(P.block_pieces(mirror, rb, cb, gr, gc, rows, cols, ro, co), mirror, rb, cb,
                     gr, gc, rows, cols, ro, co)


def test_single_rank_pieces_are_the_identity():
    for p in P.block_pieces(False, 64, 64, 1, 1, 300, 200, 17, 5):
        assert (p.local_row, p.local_col, p.rank) == (p.global_row, p.global_col, 0)


def check_ring_blocks(mirror, gr, gc, comm, m, n, ro, co, fill, tile):
    rib, cib, blocks = P.ssb_ring_blocks(mirror, gr, gc, comm, m, n, ro, co, fill, tile)
    if m == 0 or n == 0:
        assert blocks == []
        return
    assert 1 <= rib <= m and 1 <= cib <= n
    cover = np.zeros((m, n), dtype=np.int64)
    for (r, c, nr, nc) in blocks:
        assert r % rib == 0 and c % cib == 0
        assert nr == min(rib, m - r) and nc == min(cib, n - c)
        cover[r:r + nr, c:c + nc] += 1
    assert cover.max() <= 1
    # every element pgemm_ssbtr must compute lies in an active block
    i = np.arange(m)[:, None] + ro
    j = np.arange(n)[None, :] + co
    need = {0: np.ones((m, n), bool), 1: j >= i, 2: i >= j}[fill]
    assert (cover[need] == 1).all()
    # grid-wise order: the grid of each block never decreases
    grid_of = [((c // (gc * cib)), (r // (gr * rib))) for (r, c, _, _) in blocks]
    assert grid_of == sorted(grid_of)


def test_ring_blocks_pinned_by_coverage():
    rng = random.Random(11)
    for _ in range(150):
        check_ring_blocks(*random_ring_case(rng))


@pytest.mark.parametrize("case,expected", [
    # m n <= 375^2: one block
    ((0, True, 0.35, 4, 300, 400, 0, 0, 500, 375), (300, 400)),
    # mirror (not disjoint): target size
    ((0, False, 0.35, 4, 3000, 2000, 0, 0, 500, 375), (500, 500)),
    # one rank: 1 x 1 grid, factor 36 -> 6 x 6 blocks of 500
    ((0, True, 0.35, 1, 3000, 3000, 0, 0, 500, 375), (500, 500)),
    # 4 ranks: grid 2 x 2 -> 1500 x 1000, factor 6 -> (3, 2) -> 6 x 4 blocks of 500
    ((0, True, 0.35, 4, 3000, 2000, 0, 0, 500, 375), (500, 500)),
    # lower triangle: 18 of 24 blocks active, factor sqrt(4/3): 7 x 5 blocks
    ((2, True, 0.35, 4, 3000, 2000, 0, 0, 500, 375), (429, 400)),
    # blocks below the minimum: target shape stretched along the short edge
    ((0, True, 0.35, 16, 400, 3000, 0, 0, 500, 375), (400, 625)),
])
def test_block_shape_hand_cases(case, expected):
    assert P.block_size_selection(*case) == expected
    mode, disjoint, dev, comm, m, n, ro, co, tile, _ = case
    if abs(dev - 0.35) < 1e-12:
        got = spla.ssb_ring_blocks(BC if disjoint else MIRROR, 1, comm, comm, m, n, ro, co, mode,
                                   tile)
        assert got[:2] == expected


def test_find_optimal_proc_grid_hand_cases():
    assert P.find_optimal_proc_grid(1, 0, 0) == (1, 1)
    assert P.find_optimal_proc_grid(4, 1, 0) == (2, 2)
    assert P.find_optimal_proc_grid(6, 2, 0) == (2, 2)   # 4 in [4, 6], first hit
    assert P.find_optimal_proc_grid(7, 0, 0) == (7, 1)   # prime: no 2-D grid
    assert P.find_optimal_proc_grid(12, 4, 0) == (3, 3)  # 9 in [8, 12]


# ---------------------------------------------------------------------------
# planted mistakes: the pins above reject them

class _RowMajorRank(P.BlockCyclic):
    def info(self, idx):
        p = super().info(idx)
        gbr, gbc = p.global_row // self.rib, p.global_col // self.cib
        return P.Piece(p.global_row, p.global_col, p.sub_row, p.sub_col, p.local_row,
                       p.local_col, p.rows, p.cols, (gbr % self.gr) * self.gc + gbc % self.gc)


class _LocalWithoutBlockOffset(P.BlockCyclic):
    def info(self, idx):
        p = super().info(idx)
        return P.Piece(p.global_row, p.global_col, p.sub_row, p.sub_col,
                       (p.global_row // self.rib // self.gr) * self.rib, p.local_col, p.rows,
                       p.cols, p.rank)


@pytest.mark.parametrize("cls", [_RowMajorRank, _LocalWithoutBlockOffset])
def test_planted_placement_mistakes_are_rejected(cls):
    rng = random.Random(3)
    failures = 0
    for _ in range(200):
        rb, cb, gr, gc, rows, cols, ro, co = random_region(rng)
        gen = cls(rb, cb, gr, gc, rows, cols, ro, co)
        pieces = [gen.info(i) for i in range(gen.num_blocks())]
        try:
            check_pieces(pieces, False, rb, cb, gr, gc, rows, cols, ro, co)
        except AssertionError:
            failures += 1
    assert failures > 20


def test_planted_fill_mistake_is_rejected():
    # a predicate with the comparison reversed drops blocks that hold required elements
    orig = P.block_is_active
    try:
        P.block_is_active = lambda r, c, nr, nc, ro, co, mode: (
            True if mode == 0 else (c + nc + co < r + ro if mode == 1 else r + nr + ro < c + co))
        with pytest.raises(AssertionError):
            check_ring_blocks(False, 2, 2, 4, 3000, 2000, 0, 0, 2, 500)
    finally:
        P.block_is_active = orig


def test_ring_parameter_errors():
    with pytest.raises(spla.InvalidParameterError):
        spla.ssb_ring_blocks(BC, 0, 1, 1, 10, 10, 0, 0, 0, 500)
    with pytest.raises(spla.InvalidParameterError):
        spla.ssb_ring_blocks(BC, 1, 1, 1, 10, 10, -1, 0, 0, 500)
    with pytest.raises(spla.InvalidParameterError):
        spla.ssb_ring_blocks(BC, 1, 1, 1, 10, 10, 0, 0, 3, 500)
    with pytest.raises(spla.InvalidParameterError):
        spla.region_pieces(BC, 0, 4, 1, 1, 10, 10, 0, 0)
    with pytest.raises(spla.InvalidParameterError):
        spla.region_pieces(5, 4, 4, 1, 1, 10, 10, 0, 0)
