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
