"""A multi-rank pgemm_ssbtr ring, emulated on one GPU.

P virtual ranks each hold a k-stripe of A and B (SPLA's stripe distribution).
The ring plan (spla_amd_ssb_ring_blocks) gives the blocks of C in ring order;
every rank's block products are one batch (spla_amd_batch_create_ssb), and
the ring's receive-and-add (src/pgemm_ssb/ring_ssb_gpu.cpp:267-271) becomes
beta = 1 on the ranks after the first. Each finished block is placed into the
owners' local C by its pieces (spla_amd_region_pieces) with
C_local <- R + beta C_local (ring_ssb_gpu.cpp:379-384). Every rank's local C
is then compared with alpha A^H B + beta C0 at the owner / local positions of
ScaLAPACK's closed forms, so a wrong block, order, piece or product fails.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import spla_amd as spla
from helpers import make_buffer, rel_error

pytestmark = pytest.mark.gpu


def indxg2p(i, nb, nprocs):
    return (i // nb) % nprocs


def indxg2l(i, nb, nprocs):
    return (i // (nb * nprocs)) * nb + i % nb


def local_extent(glob, nb, p, nprocs):
    """Number of global indices < glob owned by process p (ScaLAPACK NUMROC)."""
    nblocks = glob // nb
    size = (nblocks // nprocs) * nb
    extra = nblocks % nprocs
    if p < extra:
        size += nb
    elif p == extra:
        size += glob % nb
    return size


@pytest.mark.parametrize("fill", [spla.FillMode.FULL, spla.FillMode.LOWER, spla.FillMode.UPPER])
def test_ring_emulation_matches_global_product(ctx, device, fill):
    import torch

    gr, gc = 2, 2
    P = gr * gc
    rb, cb = 48, 40                      # block-cyclic distribution of C
    m, n, k_rank = 600, 520, 700         # sub-matrix of C, k per rank
    ro, co = 30, 17                      # offset of the sub-matrix in the global C
    tile = 160                           # context host tile size (ring block target)
    alpha, beta = 0.75 - 0.25j, 0.5 + 0.5j
    k = k_rank * P

    rib, cib, blocks = spla.ssb_ring_blocks(0, gr, gc, P, m, n, ro, co, int(fill), tile)
    assert blocks, (rib, cib)

    A = make_buffer(21, 1, k, m, k, np.complex128)  # k x m, A^H is m x k
    B = make_buffer(21, 2, k, n, k, np.complex128)
    gm, gn = ro + m, co + n                          # global C extent
    C0 = make_buffer(21, 3, gm, gn, gm, np.complex128).reshape(gn, gm).T.copy()

    # local C of every rank, block-cyclic over the global C
    gi, gj = np.arange(gm), np.arange(gn)
    loc = {}
    for p in range(P):
        pr, pc = p % gr, p // gr
        ri, cj = gi[indxg2p(gi, rb, gr) == pr], gj[indxg2p(gj, cb, gc) == pc]
        L = np.zeros((local_extent(gm, rb, pr, gr), local_extent(gn, cb, pc, gc)),
                     dtype=np.complex128)
        assert L.shape == (len(ri), len(cj))
        L[np.ix_(indxg2l(ri, rb, gr), indxg2l(cj, cb, gc))] = C0[np.ix_(ri, cj)]
        loc[p] = L
    loc0 = {p: L.copy() for p, L in loc.items()}

    # rank stripes on the device: rank p holds rows [p k_rank, (p + 1) k_rank) of A and B
    dA = [torch.from_numpy(np.ascontiguousarray(
        A.reshape(m, k)[:, p * k_rank:(p + 1) * k_rank])).to(device) for p in range(P)]
    dB = [torch.from_numpy(np.ascontiguousarray(
        B.reshape(n, k)[:, p * k_rank:(p + 1) * k_rank])).to(device) for p in range(P)]
    R = [torch.zeros(b[2] * b[3], dtype=torch.complex128, device=device) for b in blocks]
    torch.cuda.synchronize()

    for p in range(P):  # the ring: every rank adds its share of every block
        specs = [dict(m=nr, n=nc, k=k_rank, A=dA[p].data_ptr() + 16 * r * k_rank, lda=k_rank,
                      B=dB[p].data_ptr() + 16 * c * k_rank, ldb=k_rank, C=R[i], ldc=nr)
                 for i, (r, c, nr, nc) in enumerate(blocks)]
        batch = spla.Batch(ctx, specs, alpha, 0.0 if p == 0 else 1.0)
        batch.execute()
        batch.close()

    covered = np.zeros((m, n), dtype=bool)
    for i, (r, c, nr, nc) in enumerate(blocks):
        Rb = R[i].cpu().numpy().reshape(nc, nr).T
        for q in spla.region_pieces(0, rb, cb, gr, gc, nr, nc, ro + r, co + c):
            g_r, g_c, s_r, s_c, l_r, l_c, q_r, q_c, rank = q
            L = loc[rank]
            L[l_r:l_r + q_r, l_c:l_c + q_c] = (Rb[s_r:s_r + q_r, s_c:s_c + q_c] +
                                               beta * loc0[rank][l_r:l_r + q_r, l_c:l_c + q_c])
            covered[r + s_r:r + s_r + q_r, c + s_c:c + s_c + q_c] = True

    ref = np.zeros(m * n, dtype=np.complex128)
    oracle.gemm(2, 0, m, n, k, alpha, A, k, B, k, 0.0, ref, m)
    ref = ref.reshape(n, m).T + beta * C0[ro:, co:]
    ii = np.arange(m)[:, None] + ro
    jj = np.arange(n)[None, :] + co
    need = {spla.FillMode.FULL: np.ones((m, n), bool), spla.FillMode.UPPER: jj >= ii,
            spla.FillMode.LOWER: ii >= jj}[fill]
    assert covered[need].all()
    got = np.empty((m, n), dtype=np.complex128)
    owner = indxg2p(ii, rb, gr) + indxg2p(jj, cb, gc) * gr
    li, lj = indxg2l(ii, rb, gr) + 0 * jj, indxg2l(jj, cb, gc) + 0 * ii
    for p in range(P):
        sel = owner == p
        got[sel] = loc[p][li[sel], lj[sel]]
    assert rel_error(got[need], ref[need], A, B) <= 1e-13
    # elements no block covers keep their old value, so does everything outside the sub-matrix
    assert np.array_equal(got[~covered], C0[ro:, co:][~covered])
    for p in range(P):
        pr, pc = p % gr, p // gr
        ri, cj = gi[indxg2p(gi, rb, gr) == pr], gj[indxg2p(gj, cb, gc) == pc]
        outside = (ri < ro)[:, None] | (cj < co)[None, :]
        assert np.array_equal(loc[p][outside], loc0[p][outside])


def test_sbs_ring_emulation_matches_global_product(ctx, device):
    """A multi-rank pgemm_sbs ring on one GPU (src/pgemm_sbs/pgemm_sbs_gpu.cpp:163-230,
    ring_sbs_gpu.cpp:114-368): P virtual ranks each hold a row stripe of A and C;
    B (k x n at offsets bro, bco of a block-cyclic global B) travels around the ring
    in blocks. The block shape and order are SPLA's (block_size_selection on k x n
    with full fill, grid-wise order), which spla_amd_ssb_ring_blocks computes from
    the same rules. Each block is assembled from the owners' local B by its pieces
    (spla_amd_region_pieces, as RingSBSGPU::prepare does); in a tile of P blocks
    rank p starts at block (rankOffset + p) % P and moves one block per step
    (ring), or all ranks take block s at step s (broadcast, when a tile holds
    fewer than 0.65 P blocks). Rank p's product of a step is
    C_p(:, cols) <- alpha A_p(:, rows) B_blk + beta' C_p(:, cols) with beta' = beta
    on the first visit of the column block on that rank and 1 afterwards
    (ring_sbs_gpu.cpp:278-282); the products of one step form two sbs batches (beta
    and 1). Every rank's C is compared with alpha A_p B + beta C_p0."""
    import torch

    gr, gc = 2, 2
    P = gr * gc
    rbB, cbB = 40, 48                  # block-cyclic distribution of the global B
    k, n = 900, 520                    # B sub-matrix used by the product
    bro, bco = 5, 7                    # its offset in the global B
    tile = 160                         # context host tile size (ring block target)
    mloc = [300, 0, 257, 130]          # rows of A and C on each rank
    alpha, beta = 0.5 + 0.75j, -0.25 + 0.5j

    rib, cib, blocks = spla.ssb_ring_blocks(0, gr, gc, P, k, n, bro, bco,
                                            int(spla.FillMode.FULL), tile)
    assert len(blocks) > P and rib < k  # several tiles, several visits of each column block

    gk, gn = bro + k, bco + n
    Bg = make_buffer(22, 2, gk, gn, gk, np.complex128).reshape(gn, gk).T.copy()
    gi, gj = np.arange(gk), np.arange(gn)
    locB = {}
    for p in range(P):  # each rank's local part of the block-cyclic B
        pr, pc = p % gr, p // gr
        ri, cj = gi[indxg2p(gi, rbB, gr) == pr], gj[indxg2p(gj, cbB, gc) == pc]
        L = np.zeros((local_extent(gk, rbB, pr, gr), local_extent(gn, cbB, pc, gc)),
                     dtype=np.complex128)
        L[np.ix_(indxg2l(ri, rbB, gr), indxg2l(cj, cbB, gc))] = Bg[np.ix_(ri, cj)]
        locB[p] = L

    lda = [max(1, m) + 3 for m in mloc]
    ldc = [max(1, m) + 2 for m in mloc]
    A = [make_buffer(22, 10 + p, mloc[p], k, lda[p], np.complex128) for p in range(P)]
    C0 = [make_buffer(22, 20 + p, mloc[p], n, ldc[p], np.complex128) for p in range(P)]
    dA = [torch.from_numpy(a).to(device) for a in A]
    dC = [torch.from_numpy(c.copy()).to(device) for c in C0]

    def assemble(r, c, nr, nc):
        blk = np.zeros((nr, nc), dtype=np.complex128)
        for q in spla.region_pieces(0, rbB, cbB, gr, gc, nr, nc, bro + r, bco + c):
            g_r, g_c, s_r, s_c, l_r, l_c, q_r, q_c, rank = q
            blk[s_r:s_r + q_r, s_c:s_c + q_c] = locB[rank][l_r:l_r + q_r, l_c:l_c + q_c]
        return torch.from_numpy(np.asfortranarray(blk).ravel(order="F")).to(device)

    visited = [set() for _ in range(P)]
    for t0 in range(0, len(blocks), P):
        tile_blocks = blocks[t0:t0 + P]
        r0, c0 = tile_blocks[0][0] + bro, tile_blocks[0][1] + bco
        rank_offset = indxg2p(r0, rbB, gr) + indxg2p(c0, cbB, gc) * gr
        ring = len(tile_blocks) >= P * 0.65
        dBlk = [assemble(*b) for b in tile_blocks]
        for step in range(P):
            specs = {True: [], False: []}
            for p in range(P):
                idx = (rank_offset + p + step) % P if ring else step
                if idx >= len(tile_blocks) or mloc[p] == 0:
                    continue
                r, c, nr, nc = tile_blocks[idx]
                first = c not in visited[p]
                visited[p].add(c)
                specs[first].append(dict(m=mloc[p], n=nc, k=nr,
                                         A=dA[p].data_ptr() + 16 * r * lda[p], lda=lda[p],
                                         B=dBlk[idx], ldb=nr,
                                         C=dC[p].data_ptr() + 16 * c * ldc[p], ldc=ldc[p]))
            for first, sp in specs.items():
                if sp:
                    batch = spla.Batch(ctx, sp, alpha, beta if first else 1.0,
                                       opA=spla.Operation.NONE)
                    batch.execute()
                    batch.close()

    Bsub = np.asfortranarray(Bg[bro:, bco:]).ravel(order="F")
    for p in range(P):
        got = dC[p].cpu().numpy()
        if mloc[p] == 0:
            assert np.array_equal(got, C0[p])
            continue
        assert all(c in visited[p] for _, c, _, _ in blocks)
        ref = C0[p].copy()
        oracle.gemm(0, 0, mloc[p], n, k, alpha, A[p], lda[p], Bsub, k, beta, ref, ldc[p],
                    accurate=True)
        assert rel_error(got, ref, A[p], Bsub) <= 1e-13, p
        # the ld padding of C_p is never written
        pad = got.reshape(n, ldc[p])[:, mloc[p]:]
        assert np.array_equal(pad, C0[p].reshape(n, ldc[p])[:, mloc[p]:])
