"""Parity of pgemm_ssb / pgemm_ssbtr / pgemm_sbs on one rank against the CPU oracle.

SPLA turns each of them into one local GEMM on a sub-matrix at one rank
(src/pgemm_ssb/pgemm_ssb_gpu.cpp:259-262, src/pgemm_sbs/pgemm_sbs_gpu.cpp:246-249);
SPLA's own checks are tests/test_gemm_ssb.cpp:154-523 and test_gemm_sbs.cpp:158-525
(against ScaLAPACK, 1e-6 absolute). Here: normwise error <= 1e-13 against the
double-double oracle, bit-exact placement (every entry outside the written
sub-matrix keeps its sentinel), fill modes, offsets, host-resident operands
through the k-chunked streaming pipeline, and SPLA's error behaviour.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import spla_amd as spla
from helpers import make_buffer, rel_error, to_device, to_matrix

pytestmark = pytest.mark.gpu
TOL = 1e-13


def ssb_case(ctx, device, dtype, m, n, k, row_off, col_off, fill, op="C", where="device",
             alpha=1.0, beta=0.5, chunk_bytes=None, dist=None):
    gm, gn = m + row_off + 11, n + col_off + 7  # global C with margins
    ldc = gm + 3
    lda = ldb = k + 4
    A = make_buffer(21, 1, k, m, lda, dtype)
    B = make_buffer(21, 2, k, n, ldb, dtype)
    C0 = make_buffer(21, 3, gm, gn, ldc, dtype)
    opi = 2 if op == "C" else 1
    ref = C0.copy()
    sub = row_off + col_off * ldc
    ref_sub = ref[sub:]
    oracle.gemm(opi, 0, m, n, k, alpha, A, lda, B, ldb, beta, ref_sub, ldc, accurate=True)
    ref[sub:] = ref_sub
    dist = dist or spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1,
                                                                    64, 64)
    if chunk_bytes:
        ctx.set_stream_chunk_bytes(chunk_bytes)
    try:
        if where == "device":
            dA, dB, dC = to_device(A, device), to_device(B, device), to_device(C0, device)
        else:
            dA, dB, dC = A, B, C0.copy()
        if fill is None:
            spla.pgemm_ssb(m, n, k, opi, alpha, dA, lda, dB, ldb, beta, dC, ldc, row_off, col_off,
                           dist, ctx)
        else:
            spla.pgemm_ssbtr(m, n, k, opi, alpha, dA, lda, dB, ldb, beta, dC, ldc, row_off,
                             col_off, fill, dist, ctx)
        got = dC.cpu().numpy() if where == "device" else dC
    finally:
        if chunk_bytes:
            ctx.set_stream_chunk_bytes(256 << 20)
    return A, B, C0, got, ref, (gm, gn, ldc), (lda, ldb)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("op", ["T", "C"])
@pytest.mark.parametrize("m,n,k,ro,co", [(1, 1, 1, 0, 0), (13, 32, 263, 2, 3), (263, 13, 50, 0, 5),
                                         (300, 200, 1500, 17, 9), (64, 64, 0, 1, 1)])
def test_pgemm_ssb_offsets_and_placement(ctx, device, dtype, op, m, n, k, ro, co):
    A, B, C0, got, ref, (gm, gn, ldc), (lda, ldb) = ssb_case(ctx, device, dtype, m, n, k, ro, co,
                                                             None, op=op)
    G = to_matrix(got, gm + 3, gn, ldc)
    R = to_matrix(ref, gm + 3, gn, ldc)
    mask = np.zeros(G.shape, dtype=bool)
    mask[ro:ro + m, co:co + n] = True
    assert np.array_equal(G[~mask], R[~mask])  # untouched entries bit-exact
    if k > 0:
        err = rel_error(G[mask], R[mask], to_matrix(A, k, m, lda), to_matrix(B, k, n, ldb))
        assert err <= TOL, err
    else:
        assert np.array_equal(G[mask], R[mask])  # C <- beta C


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("fill", [spla.FillMode.UPPER, spla.FillMode.LOWER, spla.FillMode.FULL])
@pytest.mark.parametrize("m,n,ro,co", [(700, 600, 0, 0), (517, 700, 129, 3), (300, 900, 5, 400)])
def test_pgemm_ssbtr_fill(ctx, device, dtype, fill, m, n, ro, co):
    k = 333
    A, B, C0, got, ref, (gm, gn, ldc), (lda, ldb) = ssb_case(ctx, device, dtype, m, n, k, ro, co,
                                                             fill)
    G = to_matrix(got, gm + 3, gn, ldc)
    R = to_matrix(ref, gm + 3, gn, ldc)
    O = to_matrix(C0, gm + 3, gn, ldc)
    rows = np.arange(G.shape[0])[:, None]
    cols = np.arange(G.shape[1])[None, :]
    inside = (rows >= ro) & (rows < ro + m) & (cols >= co) & (cols < co + n)
    if fill == spla.FillMode.UPPER:
        required = inside & (cols >= rows)
    elif fill == spla.FillMode.LOWER:
        required = inside & (rows >= cols)
    else:
        required = inside
    # required triangle computed to tolerance
    if required.any():
        err = rel_error(G[required], R[required], to_matrix(A, k, m, lda), to_matrix(B, k, n, ldb))
        assert err <= TOL, err
    # everything else: either computed or untouched (the API allows both); outside the
    # sub-matrix: untouched, bit-exact
    assert np.array_equal(G[~inside], O[~inside])
    other = inside & ~required
    close = np.abs(G[other] - R[other]) <= 1e-10 * (1 + np.abs(R[other]))
    untouched = G[other] == O[other]
    assert np.all(close | untouched)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("mloc,n,k,bro,bco", [(13, 32, 263, 0, 0), (1000, 130, 70, 5, 6),
                                              (5000, 64, 500, 0, 17)])
def test_pgemm_sbs(ctx, device, dtype, mloc, n, k, bro, bco):
    gk, gn = k + bro + 3, n + bco + 2
    ldb = gk + 1
    lda, ldc = mloc + 2, mloc + 5
    A = make_buffer(31, 1, mloc, k, lda, dtype)
    Bg = make_buffer(31, 2, gk, gn, ldb, dtype)
    C0 = make_buffer(31, 3, mloc, n, ldc, dtype)
    ref = C0.copy()
    boff = bro + bco * ldb
    oracle.gemm(0, 0, mloc, n, k, 1.5, A, lda, Bg[boff:], ldb, 0.25, ref, ldc, accurate=True)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "C", 1, 1, 128,
                                                             128)
    dC = to_device(C0, device)
    spla.pgemm_sbs(mloc, n, k, 1.5, to_device(A, device), lda, to_device(Bg, device), ldb, bro, bco,
                   dist, 0.25, dC, ldc, ctx)
    got = dC.cpu().numpy()
    err = rel_error(to_matrix(got, mloc, n, ldc), to_matrix(ref, mloc, n, ldc),
                    to_matrix(A, mloc, k, lda), to_matrix(Bg[boff:], k, n, ldb))
    assert err <= TOL, err
    pad = got.reshape(n, ldc)[:, mloc:]
    assert np.all(pad == 1234.5)


def test_pgemm_sbs_k_zero_leaves_c(ctx, device):
    # SPLA returns before touching C when k == 0 or n == 0 (pgemm_sbs_gpu.cpp:238-240).
    C0 = make_buffer(1, 3, 10, 4, 10, np.complex128)
    dC = to_device(C0, device)
    dist = spla.MatrixDistribution.create_mirror(spla.MPI_COMM_SELF)
    spla.pgemm_sbs(10, 4, 0, 1.0, None, 10, None, 1, 0, 0, dist, 0.0, dC, 10, ctx)
    assert np.array_equal(dC.cpu().numpy(), C0)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("chunk_bytes", [1 << 20, 256 << 20])
def test_host_resident_streaming(ctx, device, dtype, chunk_bytes):
    # pageable host A, B and C: k-chunks staged through pinned buffers on a copy stream
    m, n, k = 300, 257, 9000
    A, B, C0, got, ref, (gm, gn, ldc), (lda, ldb) = ssb_case(
        ctx, device, dtype, m, n, k, 3, 4, None, where="host", chunk_bytes=chunk_bytes)
    G = to_matrix(got, gm + 3, gn, ldc)
    R = to_matrix(ref, gm + 3, gn, ldc)
    err = rel_error(G[3:3 + m, 4:4 + n], R[3:3 + m, 4:4 + n], to_matrix(A, k, m, lda),
                    to_matrix(B, k, n, ldb))
    assert err <= TOL, err
    mask = np.ones(G.shape, dtype=bool)
    mask[3:3 + m, 4:4 + n] = False
    assert np.array_equal(G[mask], R[mask])


def test_pinned_host_operands(ctx, device):
    import torch

    m, n, k = 128, 96, 5000
    A = make_buffer(4, 1, k, m, k, np.complex128)
    B = make_buffer(4, 2, k, n, k, np.complex128)
    C0 = make_buffer(4, 3, m, n, m, np.complex128)
    pA = torch.from_numpy(A).pin_memory()
    pB = torch.from_numpy(B).pin_memory()
    pC = torch.from_numpy(C0.copy()).pin_memory()
    dist = spla.MatrixDistribution.create_mirror(spla.MPI_COMM_SELF)
    ctx.set_stream_chunk_bytes(1 << 20)
    try:
        spla.pgemm_ssb(m, n, k, 2, 1.0, pA, k, pB, k, 0.0, pC, m, 0, 0, dist, ctx)
    finally:
        ctx.set_stream_chunk_bytes(256 << 20)
    ref = np.zeros(m * n, dtype=np.complex128)
    oracle.gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.0, ref, m, accurate=True)
    assert rel_error(pC.numpy(), ref, A, B) <= TOL


def test_error_behaviour(ctx, device):
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 64, 64)
    buf = to_device(np.zeros(64, dtype=np.complex128), device)
    # ssb requires opA in {T, C} (pgemm_ssb_gpu.cpp:251-253)
    with pytest.raises(spla.InvalidParameterError):
        spla.pgemm_ssb(2, 2, 2, spla.Operation.NONE, 1.0, buf, 2, buf, 2, 0.0, buf, 2, 0, 0, dist,
                       ctx)
    with pytest.raises(spla.InvalidParameterError):
        spla.pgemm_ssb(2, 2, 2, 2, 1.0, buf, 2, buf, 2, 0.0, buf, 2, -1, 0, dist, ctx)
    # lda < k for opA != N (check_gemm_param.hpp:39-41)
    with pytest.raises(spla.InvalidParameterError):
        spla.pgemm_ssb(2, 2, 4, 2, 1.0, buf, 3, buf, 4, 0.0, buf, 2, 0, 0, dist, ctx)
    # null operand with non-zero sizes (check_gemm_param.hpp:46-48)
    with pytest.raises(spla.InvalidPointerError):
        spla.gemm(2, 0, 2, 2, 2, 1.0, None, 2, buf, 2, 0.0, buf, 2, ctx)
    # m == 0: no-op even with invalid other arguments (pgemm_ssb_gpu.cpp:247-249)
    spla.pgemm_ssb(0, -5, 2, spla.Operation.NONE, 1.0, None, 0, None, 0, 0.0, None, 0, -1, -1, dist,
                   ctx)
    # sbs: negative offsets (pgemm_sbs_gpu.cpp:242-244)
    with pytest.raises(spla.InvalidParameterError):
        spla.pgemm_sbs(2, 2, 2, 1.0, buf, 2, buf, 2, -1, 0, dist, 0.0, buf, 2, ctx)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32, np.complex64],
                         ids=["d", "z", "s", "c"])
@pytest.mark.parametrize("form", ["fp64", "split"])
@pytest.mark.parametrize("fill", [spla.FillMode.UPPER, spla.FillMode.LOWER])
def test_host_c_fill_beta_zero(device, dtype, form, fill):
    # Host-resident C, beta = 0 and a fill mode: C is staged to the device, the
    # tiles outside the triangle are skipped, and the whole staged C is copied
    # back. Entries outside the triangle must hold the caller's value or the
    # product (pgemm_ssbtr.h: "may or may not be computed"), never stale device
    # memory; entries outside the sub-matrix stay bit-exact.
    single = np.dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64))
    if single and form == "split":
        pytest.skip("the split form is FP64-only")
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_gemm_mode(spla.GemmMode.SPLIT_INT8 if form == "split" else spla.GemmMode.FP64)
    m, n, k, ro, co = 520, 610, 700, 9, 140
    gm, gn = m + ro + 4, n + co + 3
    ldc = gm + 2
    A = make_buffer(31, 1, k, m, k, dtype)
    B = make_buffer(31, 2, k, n, k, dtype)
    C0 = make_buffer(31, 3, gm, gn, ldc, dtype)
    # leave stale data in the device pool first (a product of the same size)
    junk = to_device(np.full(m * n, np.nan, dtype=dtype), device)
    spla.gemm(2 if np.iscomplexobj(A) else 1, 0, m, n, k, 1.0, to_device(A, device), k,
              to_device(B, device), k, 0.0, junk, m, ctx)
    del junk
    got = C0.copy()
    opi = 2 if np.iscomplexobj(A) else 1
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 64, 64)
    spla.pgemm_ssbtr(m, n, k, opi, 1.0, A, k, B, k, 0.0, got, ldc, ro, co, fill, dist, ctx)
    wide = np.complex128 if np.iscomplexobj(A) else np.float64
    ref = C0.astype(wide)
    oracle.gemm(opi, 0, m, n, k, 1.0, A.astype(wide), k, B.astype(wide), k, 0.0,
                ref[ro + co * ldc:], ldc, accurate=True)
    G = to_matrix(got, gm, gn, ldc).astype(wide)
    R = to_matrix(ref, gm, gn, ldc)
    O = to_matrix(C0, gm, gn, ldc).astype(wide)
    rows = np.arange(gm)[:, None]
    cols = np.arange(gn)[None, :]
    inside = (rows >= ro) & (rows < ro + m) & (cols >= co) & (cols < co + n)
    req = inside & ((cols >= rows) if fill == spla.FillMode.UPPER else (rows >= cols))
    tol = 1e-6 if single else TOL
    assert rel_error(G[req], R[req], A.astype(wide), B.astype(wide)) <= tol
    assert np.array_equal(G[~inside], O[~inside])
    other = inside & ~req
    assert np.isfinite(G[other]).all()
    close = np.abs(G[other] - R[other]) <= (1e-4 if single else 1e-10) * (1 + np.abs(R[other]))
    assert np.all(close | (G[other] == O[other]))
