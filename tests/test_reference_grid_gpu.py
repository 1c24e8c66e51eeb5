"""SPLA's own distributed-product test grids, run at one rank (MPI_COMM_SELF).

tests/test_gemm_ssb.cpp:477-507 and tests/test_gemm_sbs.cpp:480-510 of SPLA run
pgemm_ssb / pgemm_sbs over

* row and column block sizes {1, 64} (a 1 x 1 process grid here, so they only
  change the descriptor), and the mirror distribution;
* ssb: m, n in {1, 13, 32, 263}, the reduction length drawn from [0, 1] or
  [50, 400]; sbs: n, k in {1, 13, 32, 263}, the local rows drawn from [0, 1] or
  [50, 400];
* sub-matrix offsets (0, 0) and (2, 3) clipped to the matrix as SPLA's tests do
  (rows max(1, m - 2), offset min(m - 1, 2), test_gemm_ssb.cpp:215-218);
* alpha = 1, beta = 0, entries uniform in [0, 100), operands on the device and in
  host memory; here both the FP64 kernels and the split-precision form (forced);

and compare with ScaLAPACK's p?gemm at |C - C_ref| <= 1e-6 per component. Here the
reference is the double-double oracle of the same product; every entry outside
the written sub-matrix must be bit-identical to the input.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle
import spla_amd as spla
from helpers import to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["fp64", "split"])
def form_ctx(request, device):
    c = spla.Context(spla.ProcessingUnit.GPU)
    if request.param == "split":
        c.set_gemm_mode(spla.GemmMode.SPLIT_INT8)
    return c, request.param

SIZES = (1, 13, 32, 263)
RANGES = ((0, 1), (50, 400))
DTYPES = {"d": np.float64, "z": np.complex128}


def uniform(rng, count, dtype):
    v = rng.uniform(0.0, 100.0, size=count)
    if np.dtype(dtype) == np.complex128:
        v = v + 1j * rng.uniform(0.0, 100.0, size=count)
    return v.astype(dtype)


def distributions():
    for rb, cb in itertools.product((1, 64), (1, 64)):
        yield f"blacs{rb}x{cb}", lambda rb=rb, cb=cb: (
            spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, rb,
                                                              cb))
    yield "mirror", lambda: spla.MatrixDistribution.create_mirror(spla.MPI_COMM_SELF)


def run(fn, args, where, device):
    """Calls fn with the buffers on the device or in host memory; returns C."""
    A, B, C = args["A"], args["B"], args["C"]
    if where == "device":
        dA = to_device(A, device) if A.size else None
        dB = to_device(B, device) if B.size else None
        dC = to_device(C, device) if C.size else None
        fn(dA, dB, dC)
        return dC.cpu().numpy() if C.size else C.copy()
    got = C.copy()
    fn(A if A.size else None, B if B.size else None, got if got.size else None)
    return got


def outcome(ctx, form, reduction):
    """Split mode runs every product whose reduction has at least 50 terms in split
    precision; shorter ones may fall back because the tolerance is capped at FP64's
    own bound (reduction x 2^-53), which can need more than 53 bits per operand
    (reasons 3 and 4). FP64 mode never splits."""
    info = ctx.last_product_info()
    if form != "split":
        assert not info["split"]
    elif reduction >= 50:
        assert info["split"], info
    elif reduction > 0 and not info["split"]:
        assert info["reason"] in (3, 4), info


def check(got, ref, C0, mask):
    assert np.array_equal(got[~mask], C0[~mask])  # outside the sub-matrix: untouched
    diff = got[mask] - ref[mask]
    assert np.all(np.abs(diff.real) <= 1e-6) and np.all(np.abs(diff.imag) <= 1e-6)


@pytest.mark.parametrize("tname", list(DTYPES))
def test_pgemm_ssb_reference_grid(form_ctx, device, tname):
    ctx, form = form_ctx
    dtype = DTYPES[tname]
    rng = np.random.default_rng(4242)
    cases = 0
    for (dname, make), m, n, krange, (ro, co) in itertools.product(
            list(distributions()), SIZES, SIZES, RANGES, ((0, 0), (2, 3))):
        k = int(rng.integers(krange[0], krange[1] + 1))
        A = uniform(rng, m * k, dtype)  # k x m, ld max(k, 1)
        B = uniform(rng, n * k, dtype)  # k x n
        C0 = uniform(rng, m * n, dtype)  # m x n, ld m
        rows, cols = max(1, m - ro), max(1, n - co)
        rOff, cOff = min(m - 1, ro), min(n - 1, co)
        ld = max(k, 1)
        ref = C0.copy()
        oracle.gemm(2, 0, rows, cols, k, 1.0, A if A.size else np.zeros(1, dtype), ld,
                    B if B.size else np.zeros(1, dtype), ld, 0.0, ref[rOff + cOff * m:], m,
                    accurate=True)
        mask = np.zeros((n, m), dtype=bool)
        mask[cOff:cOff + cols, rOff:rOff + rows] = True
        mask = mask.reshape(-1)
        dist = make()
        for where in ("device", "host"):
            got = run(lambda a, b, c: spla.pgemm_ssb(
                rows, cols, k, spla.Operation.CONJ_TRANSPOSE, 1.0, a, ld, b, ld, 0.0, c, m, rOff,
                cOff, dist, ctx, dtype=dtype), dict(A=A, B=B, C=C0), where, device)
            check(got, ref, C0, mask)
            outcome(ctx, form, k)
            cases += 1
    assert cases == 5 * 16 * 2 * 2 * 2


@pytest.mark.parametrize("tname", list(DTYPES))
def test_pgemm_sbs_reference_grid(form_ctx, device, tname):
    ctx, form = form_ctx
    dtype = DTYPES[tname]
    rng = np.random.default_rng(2424)
    cases = 0
    for (dname, make), n, k, mrange, (ro, co) in itertools.product(
            list(distributions()), SIZES, SIZES, RANGES, ((0, 0), (2, 3))):
        m = int(rng.integers(mrange[0], mrange[1] + 1))
        A = uniform(rng, m * k, dtype)  # m x k, ld max(m, 1)
        B = uniform(rng, k * n, dtype)  # k x n, ld k
        C0 = uniform(rng, m * n, dtype)  # m x n, ld max(m, 1)
        rows, cols = max(1, k - ro), max(1, n - co)  # sub-matrix of B
        rOff, cOff = min(k - 1, ro), min(n - 1, co)
        lda = ldc = max(m, 1)
        ref = C0.copy()
        if m > 0:
            oracle.gemm(0, 0, m, cols, rows, 1.0, A, lda, B[rOff + cOff * k:], k, 0.0, ref, ldc,
                        accurate=True)
        mask = np.zeros((n, max(m, 1)), dtype=bool)
        mask[:cols, :m] = True
        mask = mask.reshape(-1)[: m * n]
        dist = make()
        for where in ("device", "host"):
            got = run(lambda a, b, c: spla.pgemm_sbs(
                m, cols, rows, 1.0, a, lda, b, k, rOff, cOff, dist, 0.0, c, ldc, ctx,
                dtype=dtype), dict(A=A, B=B, C=C0), where, device)
            check(got, ref, C0, mask)
            if m > 0:  # an empty product runs nothing
                outcome(ctx, form, rows)
            cases += 1
    assert cases == 5 * 16 * 2 * 2 * 2


@pytest.mark.parametrize("tname", list(DTYPES))
@pytest.mark.parametrize("form", ["ssb", "sbs"])
def test_large_reference_cases(form_ctx, device, tname, form):
    """LargeGemmSSB (m = 3000, n = 4000, reduction in [500, 6000],
    test_gemm_ssb.cpp:509-523) and LargeGemmSBS (local rows in [50, 400],
    n = 3000, k = 4000, test_gemm_sbs.cpp:514-525), offsets (2, 3), device
    operands; 512 sampled entries against the double-double oracle."""
    ctx, mode = form_ctx
    dtype = DTYPES[tname]
    rng = np.random.default_rng(77 if form == "ssb" else 78)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 128,
                                                             64)
    if form == "ssb":
        m, n, k = 3000, 4000, int(rng.integers(500, 6001))
        A, B = uniform(rng, m * k, dtype), uniform(rng, n * k, dtype)
        C0 = uniform(rng, m * n, dtype)
        rows, cols, rOff, cOff = m - 2, n - 3, 2, 3
        dC = to_device(C0, device)
        spla.pgemm_ssb(rows, cols, k, spla.Operation.CONJ_TRANSPOSE, 1.0, to_device(A, device), k,
                       to_device(B, device), k, 0.0, dC, m, rOff, cOff, dist, ctx)
        got = dC.cpu().numpy()
        ldc, opA, a, lda, b, ldb, red = m, 2, A, k, B, k, k
        off = rOff + cOff * m
    else:
        m, n, k = int(rng.integers(50, 401)), 3000, 4000
        A, B = uniform(rng, m * k, dtype), uniform(rng, k * n, dtype)
        C0 = uniform(rng, m * n, dtype)
        rows, cols, rOff, cOff = m, n - 3, 2, 3  # sub-matrix of B: (k - 2) x (n - 3) at (2, 3)
        dC = to_device(C0, device)
        spla.pgemm_sbs(m, cols, k - 2, 1.0, to_device(A, device), m, to_device(B, device), k, rOff,
                       cOff, dist, 0.0, dC, m, ctx)
        got = dC.cpu().numpy()
        ldc, opA, a, lda, b, ldb, red = m, 0, A, m, B[rOff + cOff * k:], k, k - 2
        off = 0
    mask = np.zeros((n, ldc), dtype=bool)
    r0, c0 = off % ldc, off // ldc
    mask[c0:c0 + cols, r0:r0 + rows] = True
    mask = mask.reshape(-1)
    assert np.array_equal(got[~mask], C0[~mask])
    si = rng.integers(0, rows, size=512)
    sj = rng.integers(0, cols, size=512)
    ref = oracle.gemm_entries(opA, 0, red, 1.0, a, lda, b, ldb, 0.0, None, ldc, si, sj)
    assert bool(ctx.last_product_info()["split"]) == (mode == "split")
    g = got[off + si + sj * ldc]
    diff = g - ref
    assert np.all(np.abs(diff.real) <= 1e-6) and np.all(np.abs(diff.imag) <= 1e-6), \
        np.abs(diff).max()
