"""Pins the CPU oracle (oracle/oracle.c, oracle/cpu_path.c) before it is trusted.

* exact-rational golden vectors (tests/golden/gemm_golden.json): the
  double-double oracle must reproduce the correctly rounded result (<= 1 ulp),
  the plain oracle must stay within its rounding-error bound;
* planted mistakes in the oracle's use (conjugation dropped, operand not
  transposed, beta term dropped, last reduction term dropped) must be rejected
  by the golden vectors;
* integer-valued inputs: every product / sum is exact, both restatements must
  agree bit for bit with exact integer arithmetic;
* SPLA's CPU path (host BLAS ?gemm) agrees with the oracle;
* SPLA's parameter check (src/util/check_gemm_param.hpp:36-50) on a table;
* the splitmix64 input generator: published first output and CPU/torch agreement.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "gemm_golden.json").read_text())
CASES = GOLDEN["cases"]


def dec(v, cplx):
    if cplx:
        return np.array([complex(float.fromhex(a), float.fromhex(b)) for a, b in v],
                        dtype=np.complex128)
    return np.array([float.fromhex(a) for a in v], dtype=np.float64)


def scal(v, cplx):
    return complex(float.fromhex(v[0]), float.fromhex(v[1])) if cplx else float.fromhex(v)


def load(case):
    cplx = case["type"] == "z"
    return (cplx, dec(case["A"], cplx), dec(case["B"], cplx), dec(case["C"], cplx),
            dec(case["expected"], cplx), scal(case["alpha"], cplx), scal(case["beta"], cplx))


def run(case, accurate, opA=None, opB=None, k=None, beta=None):
    cplx, A, B, C, exp, alpha, beta0 = load(case)
    out = C.copy()
    oracle.gemm(case["opA"] if opA is None else opA, case["opB"] if opB is None else opB,
                case["m"], case["n"], case["k"] if k is None else k, alpha, A, case["lda"], B,
                case["ldb"], beta0 if beta is None else beta, out, case["ldc"], accurate=accurate)
    return out, exp


def ulp_diff(a, b):
    a = np.asarray(a).view(np.float64) if np.iscomplexobj(a) else np.asarray(a)
    b = np.asarray(b).view(np.float64) if np.iscomplexobj(b) else np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    spacing = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    d = np.where(both_nan, 0.0, np.abs(a - b) / np.where(spacing > 0, spacing, 1.0))
    return float(np.max(d)) if d.size else 0.0


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_dd_oracle_matches_exact_rounding(idx):
    out, exp = run(CASES[idx], accurate=True)
    assert ulp_diff(out, exp) <= 1.0


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_plain_oracle_within_rounding_bound(idx):
    case = CASES[idx]
    out, exp = run(case, accurate=False)
    cplx, A, B, C, _, alpha, beta = load(case)
    k = case["k"]
    # |fl(sum) - sum| <= gamma_{k+2} * (|alpha| |A||B| + |beta||C|), entrywise, loosely:
    bound = (k + 4) * 2.0 ** -52 * (abs(alpha) * np.abs(A).sum() * np.abs(B).sum()
                                    + abs(beta) * np.nan_to_num(np.abs(C)).sum()) + 1e-300
    diff = np.nan_to_num(np.abs(out - exp), nan=0.0)
    assert diff.max() <= bound
    assert np.array_equal(np.isnan(out), np.isnan(exp))


def _mismatch(case, **kw):
    out, exp = run(case, accurate=True, **kw)
    return ulp_diff(out, exp) > 4.0


def _alpha_nonzero(c):
    return c["alpha"] not in ("0x0.0p+0", ["0x0.0p+0", "0x0.0p+0"])


def test_golden_vectors_reject_planted_mistakes():
    live = [c for c in CASES if _alpha_nonzero(c)]  # alpha == 0 never reads A or B
    # conjugate transpose read as plain transpose
    conj = [c for c in live if c["type"] == "z" and c["opA"] == 2 and c["k"] > 1]
    assert conj and all(_mismatch(c, opA=1) for c in conj)
    # operand not transposed (square cases only, so the call stays in bounds)
    sq = [c for c in live if c["opA"] == 1 and c["m"] == c["k"] and c["m"] > 1]
    sq += [c for c in live if c["opA"] == 0 and c["m"] == c["k"] and c["m"] > 1]
    for c in sq:
        assert _mismatch(c, opA=1 - c["opA"])
    # beta term dropped
    withc = [c for c in CASES if c["k"] > 0 and c["beta"] not in ("0x0.0p+0",
                                                                   ["0x0.0p+0", "0x0.0p+0"])]
    assert withc and all(_mismatch(c, beta=0.0) for c in withc)
    # last reduction term dropped
    longk = [c for c in live if c["k"] > 1]
    assert all(_mismatch(c, k=c["k"] - 1) for c in longk)


def test_special_cases_do_not_read_operands():
    for case in CASES:
        cplx, A, B, C, exp, alpha, beta = load(case)
        if np.isnan(C).any():  # beta == 0 with NaN in C: result finite
            out, _ = run(case, accurate=False)
            assert not np.isnan(out.reshape(case["n"], case["ldc"])[:, :case["m"]]).any()


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("opA,opB", [(0, 0), (1, 0), (2, 0), (0, 2), (2, 1)])
def test_integer_inputs_exact(cplx, opA, opB):
    rng = np.random.default_rng(5)
    if not cplx and 2 in (opA, opB):
        return
    m, n, k = 11, 9, 57
    ra, ca = (m, k) if opA == 0 else (k, m)
    rb, cb = (k, n) if opB == 0 else (n, k)
    dt = np.complex128 if cplx else np.float64

    def ints(size):
        v = rng.integers(-20, 21, size).astype(np.float64)
        if cplx:
            v = v + 1j * rng.integers(-20, 21, size)
        return v.astype(dt)

    A, B, C = ints(ra * ca), ints(rb * cb), ints(m * n)
    Am = A.reshape(ca, ra).T
    Bm = B.reshape(cb, rb).T
    opm = {0: lambda X: X, 1: lambda X: X.T, 2: lambda X: X.conj().T}
    exact = 3 * (opm[opA](Am) @ opm[opB](Bm)) - 2 * C.reshape(n, m).T
    for accurate in (False, True):
        out = C.copy()
        oracle.gemm(opA, opB, m, n, k, 3.0, A, ra, B, rb, -2.0, out, m, accurate=accurate)
        assert np.array_equal(out.reshape(n, m).T, exact)


@pytest.mark.parametrize("cplx", [False, True])
def test_cpu_path_host_blas_agrees(cplx):
    m, n, k = 37, 29, 311
    A = oracle.uniform(1, 1, k * m, cplx)
    B = oracle.uniform(1, 2, k * n, cplx)
    C = oracle.uniform(1, 3, m * n, cplx)
    ref = C.copy()
    oracle.gemm(2 if cplx else 1, 0, m, n, k, 0.5, A, k, B, k, 2.0, ref, m, accurate=True)
    got = C.copy()
    oracle.cpu_path_gemm(2 if cplx else 1, 0, m, n, k, 0.5, A, k, B, k, 2.0, got, m)
    err = oracle.normwise_error(got, ref, np.linalg.norm(A), np.linalg.norm(B))
    assert err <= 1e-15


# SPLA check_gemm_param: (opA, opB, m, n, k, A?, lda, B?, ldb, C?) -> code
PARAM_TABLE = [
    ((0, 0, 2, 3, 4, 1, 2, 1, 4, 1), 0),
    ((0, 0, -1, 3, 4, 1, 2, 1, 4, 1), 3),
    ((0, 0, 2, -3, 4, 1, 2, 1, 4, 1), 3),
    ((0, 0, 2, 3, -4, 1, 2, 1, 4, 1), 3),
    ((0, 0, 2, 3, 4, 1, 1, 1, 4, 1), 3),   # lda < m for opA = N
    ((1, 0, 2, 3, 4, 1, 3, 1, 4, 1), 3),   # lda < k for opA != N
    ((2, 0, 2, 3, 4, 1, 4, 1, 4, 1), 0),
    ((0, 0, 2, 3, 4, 1, 2, 1, 3, 1), 3),   # ldb < k for opB = N
    ((0, 1, 2, 3, 4, 1, 2, 1, 2, 1), 3),   # ldb < n for opB != N
    ((0, 2, 2, 3, 4, 1, 2, 1, 3, 1), 0),
    ((0, 0, 2, 3, 4, 0, 2, 1, 4, 1), 4),   # A null with k, m != 0
    ((0, 0, 2, 3, 4, 1, 2, 0, 4, 1), 4),
    ((0, 0, 2, 3, 4, 1, 2, 1, 4, 0), 4),
    ((0, 0, 2, 3, 0, 0, 2, 0, 0, 1), 0),   # k == 0: A, B may be null
    ((0, 0, 0, 3, 4, 0, 0, 1, 4, 0), 0),   # m == 0: A, C may be null
]


@pytest.mark.parametrize("args,code", PARAM_TABLE)
def test_check_gemm_param_table(args, code):
    assert oracle.check_gemm_param(*args) == code


def test_splitmix_generator_pins():
    # SplitMix64 (Steele, Lea, Flood 2014) seeded with 0: first output 0xE220A8397B1DCDAF
    assert int(oracle.splitmix_bits(0, 0, 0, 1)[0]) == 0xE220A8397B1DCDAF
    v = oracle.uniform(42, 1, 100000, complex_=False)
    assert v.min() >= -1.0 and v.max() < 1.0
    assert abs(v.mean()) < 0.01 and abs(v.var() - 1.0 / 3.0) < 0.01
    # the bench's device generator is the same function (checked on CPU tensors)
    torch = pytest.importorskip("torch")
    import bench

    a = bench.uniform_device(42, 2, 4099, True, torch.device("cpu"), chunk=1000).numpy()
    assert np.array_equal(a, oracle.uniform(42, 2, 4099, complex_=True))
