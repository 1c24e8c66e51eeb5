"""Pins the two oracle functions that check the largest GPU results.

``oracle.gemm_entries`` (oracle/oracle.c, oracle_gemm_{d,z}_dd_entries) is the
checker of the 3000 x 4000 products of SPLA's own test grids; ``ozaki.emulate``
(oracle/ozaki.py) is the bit-exact checker of the split-precision form.

gemm_entries:
* every exact-rational golden vector (tests/golden/gemm_golden.json, all
  operand combinations, alpha / beta != 0, NaN in C with beta = 0, k = 0):
  every entry of the result, requested one (row, col) pair at a time, within
  1 ulp of the correctly rounded value;
* planted mistakes rejected: rows and columns swapped, a wrong leading
  dimension of C, the conjugation dropped.

ozaki.emulate (alpha = 1, beta = 0):
* exact rationals: entries q 2^s with |q| < 2^P and one shift per row / column,
  so every scaled entry is an integer and the scaled product fits in 53 bits;
  the result must then equal the exact rational product (Python Fractions of
  the inputs) bit for bit, for N / T / C on either side, real and complex, with
  padded leading dimensions;
* the limb conversion of a large integer is within 2 ulp of the correctly
  rounded double (exact below 2^53);
* against the double-double oracle on exponent-spread data (entries
  U[-1, 1] 2^e, e in [-10, 10]) at the widths plan() chooses: normwise error
  within the plan's bound;
* planted mistakes rejected: conjugation dropped, exponent off by one (one
  bit less than P), operand read untransposed, last reduction term dropped.
"""

from __future__ import annotations

import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import ozaki

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "gemm_golden.json").read_text())
CASES = GOLDEN["cases"]


def dec(v, cplx):
    if cplx:
        return np.array([complex(float.fromhex(a), float.fromhex(b)) for a, b in v],
                        dtype=np.complex128)
    return np.array([float.fromhex(a) for a in v], dtype=np.float64)


def scal(v, cplx):
    return complex(float.fromhex(v[0]), float.fromhex(v[1])) if cplx else float.fromhex(v)


def ulp_diff(a, b):
    a = np.asarray(a).view(np.float64) if np.iscomplexobj(a) else np.asarray(a, dtype=np.float64)
    b = np.asarray(b).view(np.float64) if np.iscomplexobj(b) else np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    spacing = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    d = np.where(both_nan, 0.0, np.abs(a - b) / np.where(spacing > 0, spacing, 1.0))
    return float(np.max(d)) if d.size else 0.0


def entries_of(case, rows=None, cols=None, ldc=None, opA=None):
    cplx = case["type"] == "z"
    A, B, C = dec(case["A"], cplx), dec(case["B"], cplx), dec(case["C"], cplx)
    m, n = case["m"], case["n"]
    if rows is None:
        jj, ii = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
        rows, cols = ii.ravel(), jj.ravel()
    out = oracle.gemm_entries(case["opA"] if opA is None else opA, case["opB"], case["k"],
                              scal(case["alpha"], cplx), A, case["lda"], B, case["ldb"],
                              scal(case["beta"], cplx), C, case["ldc"] if ldc is None else ldc,
                              rows, cols)
    exp = dec(case["expected"], cplx)
    return out, exp[np.asarray(rows) + np.asarray(cols) * case["ldc"]]


LIVE = [c for c in CASES if c["m"] > 0 and c["n"] > 0]


@pytest.mark.parametrize("idx", range(len(LIVE)))
def test_gemm_entries_match_golden(idx):
    out, exp = entries_of(LIVE[idx])
    assert ulp_diff(out, exp) <= 1.0


def _zero(v):
    return v in ("0x0.0p+0", ["0x0.0p+0", "0x0.0p+0"])


def test_gemm_entries_reject_planted_mistakes():
    # rows and columns swapped, on the leading square part so the call stays in bounds
    sq = [c for c in LIVE if min(c["m"], c["n"]) > 1 and not _zero(c["alpha"])]
    assert sq
    for c in sq:
        s = min(c["m"], c["n"])
        jj, ii = np.meshgrid(np.arange(s), np.arange(s), indexing="ij")
        _, exp = entries_of(c, rows=ii.ravel(), cols=jj.ravel())
        out, _ = entries_of(c, rows=jj.ravel(), cols=ii.ravel())
        assert ulp_diff(out, exp) > 4.0, c
    # a wrong leading dimension of C (beta != 0, more than one column, still in bounds)
    withc = [c for c in LIVE if not _zero(c["beta"]) and c["n"] > 1 and c["ldc"] > c["m"]
             and not np.isnan(dec(c["C"], c["type"] == "z")).any()]
    assert withc
    for c in withc:
        jj, ii = np.meshgrid(np.arange(c["n"]), np.arange(c["m"]), indexing="ij")
        out, exp = entries_of(c, rows=ii.ravel(), cols=jj.ravel(), ldc=c["ldc"] - 1)
        assert ulp_diff(out, exp) > 4.0, c
    # conjugation dropped
    conj = [c for c in LIVE if c["type"] == "z" and c["opA"] == 2 and c["k"] > 0
            and not _zero(c["alpha"])]
    assert conj and all(ulp_diff(*entries_of(c, opA=1)) > 4.0 for c in conj)


# ---------------------------------------------------------------------------
# ozaki.emulate


def _dyadic(rng, count, P, shift, cplx):
    """Entries q 2^shift with integer |q| < 2^P (one shift for the whole vector)."""
    q = rng.integers(-(2 ** P) + 1, 2 ** P, count)
    v = np.ldexp(q.astype(np.float64), shift)
    if cplx:
        v = v + 1j * np.ldexp(rng.integers(-(2 ** P) + 1, 2 ** P, count).astype(np.float64), shift)
    return v


def _stored(rng, op, rows, cols, ld, P, cplx, by_row):
    """Column-major buffer of X, where op(X) is rows x cols (reduction along cols for A,
    rows for B). by_row: one exponent shift per reduction vector of op(X)."""
    st_r, st_c = (rows, cols) if op == 0 else (cols, rows)
    buf = np.full(max(1, ld * st_c), 777.0, dtype=np.complex128 if cplx else np.float64)
    M = np.zeros((st_r, st_c), dtype=buf.dtype)
    vecs = rows if by_row else cols
    length = cols if by_row else rows
    for v in range(vecs):
        x = _dyadic(rng, length, P, int(rng.integers(-30, 30)), cplx)
        # vector v of op(X): a row of op(X) (A) or a column of op(X) (B)
        if by_row:
            if op == 0:
                M[v, :] = x
            else:
                M[:, v] = np.conj(x) if op == 2 else x
        else:
            if op == 0:
                M[:, v] = x
            else:
                M[v, :] = np.conj(x) if op == 2 else x
    buf[: ld * st_c].reshape(st_c, ld)[:, :st_r] = M.T
    return buf, M


def _exact_product(opA, opB, MA, MB, cplx):
    """alpha = 1, beta = 0 product of the stored matrices in Fractions, as doubles."""
    def opm(M, op):
        return M if op == 0 else (M.T.conj() if op == 2 else M.T)

    a, b = opm(MA, opA), opm(MB, opB)
    m, k = a.shape
    n = b.shape[1]

    def fr(x):
        return Fraction(float(x))

    out = np.empty((m, n), dtype=np.complex128 if cplx else np.float64)
    for i in range(m):
        for j in range(n):
            if cplx:
                re = sum(fr(a[i, l].real) * fr(b[l, j].real) - fr(a[i, l].imag) * fr(b[l, j].imag)
                         for l in range(k))
                im = sum(fr(a[i, l].real) * fr(b[l, j].imag) + fr(a[i, l].imag) * fr(b[l, j].real)
                         for l in range(k))
                # exact in double: |numerator| < 2^53 by construction
                assert float(re) == re and float(im) == im
                out[i, j] = complex(float(re), float(im))
            else:
                v = sum(fr(a[i, l]) * fr(b[l, j]) for l in range(k))
                assert float(v) == v
                out[i, j] = float(v)
    return out


OPS = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 2), (2, 2)]


def _exact_case(seed, cplx, opA, opB, P=12):
    rng = np.random.default_rng(seed)
    m, n, k = 5, 4, 23
    lda = (m if opA == 0 else k) + 2
    ldb = (k if opB == 0 else n) + 3
    A, MA = _stored(rng, opA, m, k, lda, P, cplx, by_row=True)
    B, MB = _stored(rng, opB, k, n, ldb, P, cplx, by_row=False)
    return (m, n, k, A, lda, B, ldb), MA, MB


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("opA,opB", OPS)
def test_emulate_matches_exact_rationals(cplx, opA, opB):
    if not cplx and 2 in (opA, opB):
        pytest.skip("conjugation is the identity on real data")
    P = 12
    (m, n, k, A, lda, B, ldb), MA, MB = _exact_case(3 + 7 * opA + opB, cplx, opA, opB, P)
    exact = _exact_product(opA, opB, MA, MB, cplx)
    got = ozaki.emulate(opA, opB, m, n, k, A, lda, B, ldb, P, P)
    assert np.array_equal(got, exact)


def test_emulate_exact_rationals_reject_planted_mistakes(monkeypatch):
    P = 12
    # conjugation dropped: op_rows without np.conj
    (m, n, k, A, lda, B, ldb), MA, MB = _exact_case(17, True, 2, 0, P)
    exact = _exact_product(2, 0, MA, MB, True)
    real_op_rows = ozaki.op_rows

    def no_conj(X, ld, op, rows, k_):
        return real_op_rows(X, ld, 1 if op == 2 else op, rows, k_)

    monkeypatch.setattr(ozaki, "op_rows", no_conj)
    assert not np.array_equal(ozaki.emulate(2, 0, m, n, k, A, lda, B, ldb, P, P), exact)
    monkeypatch.undo()
    assert np.array_equal(ozaki.emulate(2, 0, m, n, k, A, lda, B, ldb, P, P), exact)

    # exponent off by one: one bit less than P per entry (rounds odd integers)
    (m, n, k, A, lda, B, ldb), MA, MB = _exact_case(18, False, 1, 0, P)
    exact = _exact_product(1, 0, MA, MB, False)
    monkeypatch.setattr(ozaki, "ilogb", lambda x: math.frexp(x)[1])
    assert not np.array_equal(ozaki.emulate(1, 0, m, n, k, A, lda, B, ldb, P, P), exact)
    monkeypatch.undo()

    # operand read untransposed: square op(A) read as N instead of T
    rng = np.random.default_rng(19)
    k = m = 6
    A, MA = _stored(rng, 1, m, k, k, P, False, by_row=True)
    B, MB = _stored(rng, 0, k, 4, k, P, False, by_row=False)
    exact = _exact_product(1, 0, MA, MB, False)
    assert np.array_equal(ozaki.emulate(1, 0, m, 4, k, A, k, B, k, P, P), exact)
    assert not np.array_equal(ozaki.emulate(0, 0, m, 4, k, A, k, B, k, P, P), exact)

    # last reduction term dropped
    (m, n, k, A, lda, B, ldb), MA, MB = _exact_case(20, True, 2, 0, P)
    exact = _exact_product(2, 0, MA, MB, True)
    assert not np.array_equal(ozaki.emulate(2, 0, m, n, k - 1, A, lda, B, ldb, P, P), exact)


def test_limb_conversion_close_to_correct_rounding():
    rng = np.random.default_rng(4)
    for bits in (10, 52, 53, 54, 64, 65, 100, 127, 128, 129, 150, 190):
        for _ in range(200):
            v = int(rng.integers(0, 2 ** 62)) << max(0, bits - 62)
            v ^= int(rng.integers(0, 2 ** 31))
            v = v if rng.integers(0, 2) else -v
            d = ozaki.limb_to_double(v)
            ref = float(v)  # correctly rounded
            if abs(v) < 2 ** 53:
                assert d == ref
            else:
                assert abs(d - ref) <= 2 * np.spacing(abs(ref)), (bits, v)


@pytest.mark.parametrize("cplx", [False, True])
def test_emulate_within_plan_bound_against_dd_oracle(cplx):
    rng = np.random.default_rng(30 + cplx)
    m, n, k = 12, 10, 300
    comps = 2 if cplx else 1
    opA = 2 if cplx else 1

    def spread(count):
        v = rng.uniform(-1, 1, count) * np.exp2(rng.integers(-10, 11, count))
        if cplx:
            v = v + 1j * rng.uniform(-1, 1, count) * np.exp2(rng.integers(-10, 11, count))
        return v

    A, B = spread(k * m), spread(k * n)
    opa = ozaki.op_rows(A, k, opA, m, k)
    opb = B.reshape(n, k)

    def stats(rows):
        c = np.concatenate([rows.real, rows.imag], axis=1) if cplx else rows
        return np.max(np.abs(c), axis=1), np.sum(np.abs(c) ** 2, axis=1)

    mxA, sqA = stats(opa)
    mxB, sqB = stats(opb)
    for tol in (1e-6, 1e-10, 1e-13):
        p = ozaki.plan(mxA, sqA, mxB, sqB, k, comps, tol)
        assert p["ok"], p
        got = ozaki.emulate(opA, 0, m, n, k, A, k, B, k, p["PA"], p["PB"])
        ref = np.zeros(m * n, dtype=A.dtype)
        oracle.gemm(opA, 0, m, n, k, 1.0, A, k, B, k, 0.0, ref, m, accurate=True)
        err = oracle.normwise_error(got.T.ravel(), ref, np.linalg.norm(A), np.linalg.norm(B))
        assert err <= p["bound"], (tol, err, p["bound"])
        # four bits fewer than planned would break the bound's premise: the
        # restatement at P - 4 is measurably worse than at P (the bound is sharp
        # enough to separate them by at least a factor 4 in the observed error)
        worse = ozaki.emulate(opA, 0, m, n, k, A, k, B, k, p["PA"] - 4, p["PB"] - 4)
        err4 = oracle.normwise_error(worse.T.ravel(), ref, np.linalg.norm(A), np.linalg.norm(B))
        assert err4 > 4 * err
