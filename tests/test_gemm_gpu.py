"""Parity of spla::gemm on the GPU against the CPU oracle.

Re-expresses SPLA's tests/test_gemm.cpp:35-254 (m, n, k in {1, 13, 32, 263};
opA, opB in {N, C}; alpha = 2, beta = 3; lda = rows + 5, ldb = rows + 6,
ldc = m + 7; host and device operands) with complex-valued data (SPLA's test
only fills real parts), plus the transpose operation, BLAS special cases,
integer-valued bit-exact checks, split-K sizes and run-to-run determinism.
Tolerance: normwise ||C - C_dd||_F / (||A||_F ||B||_F) <= 1e-13 against the
double-double oracle; integer-valued inputs must match bit for bit.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle
import spla_amd as spla
from helpers import OPS, make_buffer, op_shape, rel_error, to_device, to_matrix

pytestmark = pytest.mark.gpu

SIZES = (1, 13, 32, 263)
TOL = 1e-13


def run_case(ctx, device, dtype, opA, opB, m, n, k, alpha=2.0, beta=3.0, where="device",
             kind="uniform", seed=7):
    oa, ob = OPS[opA], OPS[opB]
    ra, ca = op_shape(oa, m, k)
    rb, cb = op_shape(ob, k, n)
    lda, ldb, ldc = ra + 5, rb + 6, m + 7
    A = make_buffer(seed, 1, ra, ca, lda, dtype, kind)
    B = make_buffer(seed, 2, rb, cb, ldb, dtype, kind)
    C0 = make_buffer(seed, 3, m, n, ldc, dtype, kind)
    ref = C0.copy()
    oracle.gemm(oa, ob, m, n, k, alpha, A, lda, B, ldb, beta, ref, ldc, accurate=True)
    if where == "device":
        dA, dB, dC = to_device(A, device), to_device(B, device), to_device(C0, device)
        spla.gemm(oa, ob, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, ctx)
        got = dC.cpu().numpy()
    else:
        got = C0.copy()
        spla.gemm(oa, ob, m, n, k, alpha, A, lda, B, ldb, beta, got, ldc, ctx)
    return A, B, C0, got, ref, (ra, ca, lda), (rb, cb, ldb), ldc


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("opA,opB", list(itertools.product("NC", "NC")) + [("T", "N"), ("N", "T"),
                                                                           ("T", "T"), ("C", "T")])
@pytest.mark.parametrize("where", ["device", "host"])
def test_gemm_reference_grid(ctx, device, dtype, opA, opB, where):
    worst = 0.0
    for m, n, k in itertools.product(SIZES, SIZES, SIZES):
        A, B, C0, got, ref, sa, sb, ldc = run_case(ctx, device, dtype, opA, opB, m, n, k,
                                                   where=where)
        err = rel_error(to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc),
                        to_matrix(A, *sa), to_matrix(B, *sb))
        worst = max(worst, err)
        assert err <= TOL, (m, n, k, err)
        # ld padding of C untouched
        pad = got[: ldc * n].reshape(n, ldc)[:, m:]
        assert np.all(pad == 1234.5), (m, n, k)
    print(f"worst normwise error {worst:.2e}")


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("opA,opB", [("C", "N"), ("N", "N"), ("T", "C"), ("N", "C")])
def test_gemm_integer_inputs_bit_exact(ctx, device, dtype, opA, opB):
    # |entries| <= 8 and k <= 263: every product and partial sum is an exact
    # FP64 integer, so any summation order gives the same bits as the oracle.
    for m, n, k in [(13, 32, 263), (263, 13, 32), (130, 70, 263)]:
        _, _, _, got, ref, _, _, _ = run_case(ctx, device, dtype, opA, opB, m, n, k, kind="int")
        assert np.array_equal(got, ref), (m, n, k)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("tile", ["L", "XL"])
@pytest.mark.parametrize("opA,opB", [("C", "N"), ("N", "N"), ("T", "T"), ("N", "T")])
def test_large_tiles_every_kblock_count_bit_exact(ctx, device, dtype, tile, opA, opB,
                                                  monkeypatch):
    """The 128 x 128 (L) and 256 x 128 (XL) tiles, forced, on products of 1 to 5 k-blocks
    of 16 embedded reduction elements in one and in two splits: the first and last
    k-blocks of each schedule (stores and reads around the barrier, the loads of the
    k-block after next), every operand layout (K-major or MN-major A and B) and edge
    tiles in m and n. Integer inputs: bit-exact against the oracle."""
    monkeypatch.setenv("SPLA_AMD_TILE", tile)
    per_kblock = 8 if dtype == np.complex128 else 16
    for splits in ("1", "2"):
        monkeypatch.setenv("SPLA_AMD_SPLITS", splits)
        for kb in (1, 2, 3, 4, 5):
            for m, n in ((130, 140), (64, 128)):
                k = kb * per_kblock
                _, _, _, got, ref, _, _, _ = run_case(ctx, device, dtype, opA, opB, m, n, k,
                                                      kind="int")
                assert np.array_equal(got, ref), (splits, kb, m, n)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
def test_beta_zero_does_not_read_c(ctx, device, dtype):
    m, n, k = 40, 33, 70
    A = make_buffer(3, 1, k, m, k, dtype)
    B = make_buffer(3, 2, k, n, k, dtype)
    C = np.full(m * n, np.nan, dtype=dtype)
    dC = to_device(C, device)
    spla.gemm(OPS["C"], OPS["N"], m, n, k, 1.5, to_device(A, device), k, to_device(B, device), k,
              0.0, dC, m, ctx)
    got = dC.cpu().numpy()
    ref = np.zeros(m * n, dtype=dtype)
    oracle.gemm(2, 0, m, n, k, 1.5, A, k, B, k, 0.0, ref, m, accurate=True)
    assert np.all(np.isfinite(got))
    assert rel_error(got, ref, A, B) <= TOL


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
def test_alpha_zero_and_k_zero_scale_c(ctx, device, dtype):
    m, n, k = 17, 9, 5
    A = np.full(k * m, np.nan, dtype=dtype)
    B = np.full(k * n, np.nan, dtype=dtype)
    C0 = make_buffer(5, 3, m, n, m, dtype)
    for kk, alpha in ((k, 0.0), (0, 2.0)):
        dC = to_device(C0, device)
        spla.gemm(OPS["T"], OPS["N"], m, n, kk, alpha, to_device(A, device), max(kk, 1),
                  to_device(B, device), max(kk, 1), 3.0, dC, m, ctx)
        assert np.array_equal(dC.cpu().numpy(), 3.0 * C0)
    # beta == 0 as well: zeros, C not read
    dC = to_device(np.full(m * n, np.nan, dtype=dtype), device)
    spla.gemm(OPS["T"], OPS["N"], m, n, 0, 2.0, None, 1, None, 1, 0.0, dC, m, ctx)
    assert np.all(dC.cpu().numpy() == 0)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("shape", [(64, 48, 20000), (200, 300, 7000), (1000, 17, 3000),
                                   (5, 700, 9000)])
def test_split_k_shapes(ctx, device, dtype, shape):
    m, n, k = shape
    alpha = (1.0 - 0.5j) if dtype == np.complex128 else -1.25
    A, B, _, got, ref, sa, sb, ldc = run_case(ctx, device, dtype, "C", "N", m, n, k,
                                              alpha=alpha, beta=0.5)
    err = rel_error(to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc), to_matrix(A, *sa),
                    to_matrix(B, *sb))
    assert err <= TOL, err


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("shape", [(16, 16, 400000), (1, 1, 1000000), (33, 7, 300000)])
def test_small_mn_long_reduction(ctx, device, dtype, shape):
    """HBM-bound shapes: hundreds of split-K partial sums, reduced by the wide
    fixed-order reduction; accurate and bitwise identical from run to run."""
    import torch

    m, n, k = shape
    A, B, C0, got, ref, sa, sb, ldc = run_case(ctx, device, dtype, "C", "N", m, n, k,
                                               alpha=1.5, beta=0.5)
    err = rel_error(to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc), to_matrix(A, *sa),
                    to_matrix(B, *sb))
    assert err <= TOL, err
    dA, dB = to_device(A, device), to_device(B, device)
    outs = []
    for _ in range(2):
        dC = to_device(C0, device)
        spla.gemm(spla.Operation.CONJ_TRANSPOSE, spla.Operation.NONE, m, n, k, 1.5, dA, sa[2], dB,
                  sb[2], 0.5, dC, ldc, ctx)
        torch.cuda.synchronize()
        outs.append(dC.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


# the streaming kernel for small m, n (K-major A and B, long reductions): every
# fragment-count variant (complex rows per fragment 8, real 16; columns 16), partial
# last steps (k not a multiple of the 8-element sub-block), rows / columns past m / n
SKINNY = {np.complex128: [(8, 16, 100003), (5, 30, 50001), (16, 9, 77777), (20, 17, 40000),
                          (32, 32, 60000), (1, 1, 9000)],
          np.float64: [(16, 16, 100001), (3, 31, 50000), (30, 12, 77779), (64, 32, 30001),
                       (40, 3, 20000)]}


@pytest.mark.parametrize("dtype", [np.complex128, np.float64], ids=["z", "d"])
@pytest.mark.parametrize("op", ["C", "T"])
def test_skinny_streaming_shapes(ctx, device, dtype, op):
    if dtype == np.float64 and op == "C":
        pytest.skip("conjugation is the identity for real operands")
    for m, n, k in SKINNY[dtype]:
        A, B, C0, got, ref, sa, sb, ldc = run_case(ctx, device, dtype, op, "N", m, n, k,
                                                   alpha=(0.5 - 2j) if dtype == np.complex128
                                                   else 1.5, beta=0.25)
        err = rel_error(to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc), to_matrix(A, *sa),
                        to_matrix(B, *sb))
        assert err <= TOL, (m, n, k, err)
        pad = got[: ldc * n].reshape(n, ldc)[:, m:]
        assert np.all(pad == 1234.5), (m, n, k)
        # integer-valued inputs: every partial sum exact, so the result is bit-exact
        _, _, _, gi, ri, _, _, _ = run_case(ctx, device, dtype, op, "N", m, n, min(k, 40000),
                                            kind="int", alpha=2.0, beta=-1.0)
        assert np.array_equal(gi, ri), (m, n, k)


def test_run_to_run_bitwise_deterministic(ctx, device):
    m, n, k = 96, 80, 30000  # split-K active
    A = make_buffer(9, 1, k, m, k, np.complex128)
    B = make_buffer(9, 2, k, n, k, np.complex128)
    dA, dB = to_device(A, device), to_device(B, device)
    outs = []
    for _ in range(3):
        dC = to_device(np.zeros(m * n, dtype=np.complex128), device)
        spla.gemm(OPS["C"], OPS["N"], m, n, k, 1.0, dA, k, dB, k, 0.0, dC, m, ctx)
        outs.append(dC.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("dtype", [np.float32, np.complex64], ids=["s", "c"])
def test_single_precision_entry_points(ctx, device, dtype):
    for opA, opB in (("N", "N"), ("C", "N"), ("T", "C")):
        for m, n, k in ((13, 32, 263), (263, 1, 13)):
            oa, ob = OPS[opA], OPS[opB]
            ra, ca = op_shape(oa, m, k)
            rb, cb = op_shape(ob, k, n)
            A = make_buffer(11, 1, ra, ca, ra + 1, dtype)
            B = make_buffer(11, 2, rb, cb, rb + 2, dtype)
            C0 = make_buffer(11, 3, m, n, m + 3, dtype)
            wide = np.complex128 if np.iscomplexobj(C0) else np.float64
            ref = C0.astype(wide)
            oracle.gemm(oa, ob, m, n, k, 2.0, A.astype(wide), ra + 1, B.astype(wide), rb + 2, 3.0,
                        ref, m + 3, accurate=True)
            for where in ("device", "host"):
                if where == "device":
                    dC = to_device(C0, device)
                    spla.gemm(oa, ob, m, n, k, 2.0, to_device(A, device), ra + 1,
                              to_device(B, device), rb + 2, 3.0, dC, m + 3, ctx)
                    got = dC.cpu().numpy()
                else:
                    got = C0.copy()
                    spla.gemm(oa, ob, m, n, k, 2.0, A, ra + 1, B, rb + 2, 3.0, got, m + 3, ctx)
                err = rel_error(to_matrix(got, m, n, m + 3), to_matrix(ref, m, n, m + 3),
                                to_matrix(A, ra, ca, ra + 1), to_matrix(B, rb, cb, rb + 2))
                assert err <= 2e-7, (opA, opB, m, n, k, where, err)
