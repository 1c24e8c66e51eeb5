"""Host tests of the split-precision planner (spla_amd/csrc/runtime/split_plan.cpp).

* the CRT constants: W_t = 1 (mod m_t), W_t = 0 (mod m_s) for s != t, W_t < M,
  M = prod m_t (checked with Python integers);
* the moduli are pairwise coprime and fit INT8 symmetric residues;
* the C++ plan equals the restatement in oracle/ozaki.py on random statistics;
* the plan's bound is an upper bound of the rounding error of the scaled
  operands it prescribes (brute force on small random matrices), the bound is
  below the tolerance, and the moduli cover 2 max |a^T b|;
* fall-back reasons: non-finite input, zero operand, exponent range too wide,
  too many moduli.
"""

import ctypes
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import ozaki

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@pytest.fixture(scope="module")
def planner(tmp_path_factory):
    out = tmp_path_factory.mktemp("plan") / "libplan.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC",
                    "-I", str(ROOT / "spla_amd" / "csrc"),
                    str(ROOT / "spla_amd" / "csrc" / "runtime" / "split_plan.cpp"),
                    str(ROOT / "tests" / "cpp" / "split_plan_capi.cpp"), "-o", str(out)],
                   check=True)
    lib = ctypes.CDLL(str(out))
    d, i = ctypes.c_void_p, ctypes.c_int
    lib.plan_c.argtypes = [d, d, i, d, d, i, ctypes.c_longlong, i, ctypes.c_double, i, d, d, d, d]
    lib.crt_c.argtypes = [i, d, d, d, d]
    lib.min_moduli_c.argtypes = [ctypes.c_double]
    lib.i8_shape_c.argtypes = [ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, i, i, d]
    lib.i8_shape_c.restype = ctypes.c_double
    return lib


def run_plan(lib, mA, sA, mB, sB, k, comps, tol, max_moduli=20):
    mA, sA, mB, sB = (np.ascontiguousarray(x, dtype=np.float64) for x in (mA, sA, mB, sB))
    out = np.zeros(5, dtype=np.int32)
    bound = ctypes.c_double()
    eA = np.zeros(len(mA), dtype=np.int32)
    eB = np.zeros(len(mB), dtype=np.int32)
    lib.plan_c(mA.ctypes.data, sA.ctypes.data, len(mA), mB.ctypes.data, sB.ctypes.data, len(mB),
               k, comps, tol, max_moduli, out.ctypes.data, ctypes.byref(bound), eA.ctypes.data,
               eB.ctypes.data)
    return dict(ok=bool(out[0]), reason=int(out[1]), PA=int(out[2]), PB=int(out[3]),
                moduli=int(out[4]), bound=bound.value, expA=list(eA), expB=list(eB))


def test_moduli_pairwise_coprime():
    for i, a in enumerate(ozaki.MODULI):
        assert 2 <= a <= 256
        for b in ozaki.MODULI[i + 1:]:
            assert math.gcd(a, b) == 1


@pytest.mark.parametrize("count", [1, 2, 7, 14, 20])
def test_crt_constants(planner, count):
    W = np.zeros(3 * count, dtype=np.uint64)
    M = np.zeros(3, dtype=np.uint64)
    H = np.zeros(3, dtype=np.uint64)
    Md = ctypes.c_double()
    planner.crt_c(count, W.ctypes.data, M.ctypes.data, H.ctypes.data, ctypes.byref(Md))
    limbs = lambda v: int(v[0]) + (int(v[1]) << 64) + (int(v[2]) << 128)  # noqa: E731
    Wp, Mp = ozaki.crt_constants(count)
    assert limbs(M) == Mp == math.prod(ozaki.MODULI[:count])
    assert limbs(H) == Mp // 2
    assert Md.value == float(Mp)
    for t in range(count):
        w = limbs(W[3 * t:3 * t + 3])
        assert w == Wp[t] and w < Mp
        for s, m in enumerate(ozaki.MODULI[:count]):
            assert w % m == (1 if s == t else 0)


def random_stats(rng, rows, k, comps, spread):
    e = rng.integers(-spread, spread + 1, size=rows)
    mx = rng.uniform(0.5, 1.0, size=rows) * np.exp2(e)
    ss = mx ** 2 * k * comps * rng.uniform(0.2, 0.6, size=rows)
    return mx, ss


@pytest.mark.parametrize("seed", range(8))
def test_plan_matches_restatement(planner, seed):
    rng = np.random.default_rng(seed)
    comps = 1 + seed % 2
    k = int(rng.integers(10, 300000))
    mA, sA = random_stats(rng, int(rng.integers(1, 300)), k, comps, int(rng.integers(0, 30)))
    mB, sB = random_stats(rng, int(rng.integers(1, 300)), k, comps, int(rng.integers(0, 30)))
    tol = float(10.0 ** rng.uniform(-15, -11))
    got = run_plan(planner, mA, sA, mB, sB, k, comps, tol)
    ref = ozaki.plan(mA, sA, mB, sB, k, comps, tol)
    for key in ("ok", "reason", "PA", "PB", "moduli"):
        assert got[key] == ref[key], key
    assert got["bound"] == pytest.approx(ref["bound"], rel=1e-12)
    if got["ok"]:
        assert got["expA"] == ref["expA"] and got["expB"] == ref["expB"]
        assert got["bound"] <= tol


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("spread", [0, 10])
def test_plan_bound_and_moduli_cover_product(planner, cplx, spread):
    rng = np.random.default_rng(11 + spread)
    m, n, k = 12, 9, 257
    comps = 2 if cplx else 1

    def mat(rows):
        e = rng.integers(-spread, spread + 1, size=(rows, k))
        v = rng.uniform(-1, 1, size=(rows, k)) * np.exp2(e)
        if cplx:
            v = v + 1j * rng.uniform(-1, 1, size=(rows, k)) * np.exp2(e)
        return v

    a, b = mat(m), mat(n)  # rows of op(A), columns of op(B)

    def stats(x):
        comp = np.concatenate([x.real, x.imag], axis=1) if cplx else x
        return np.abs(comp).max(axis=1), (np.abs(x) ** 2).sum(axis=1)

    tol = 1e-13
    p = run_plan(planner, *stats(a), *stats(b), k, comps, tol)
    assert p["ok"] and p["bound"] <= tol

    def scaled(x, e):
        f = np.exp2(np.asarray(e, dtype=np.float64))[:, None]
        return np.rint(x.real * f) + (1j * np.rint(x.imag * f) if cplx else 0), f

    ia, fa = scaled(a, p["expA"])
    ib, fb = scaled(b, p["expB"])
    assert np.abs(np.concatenate([ia.real, ia.imag] if cplx else [ia], axis=1)).max() <= 2.0 ** p["PA"]
    rhoA = np.linalg.norm(ia / fa - a) / np.linalg.norm(a)
    rhoB = np.linalg.norm(ib / fb - b) / np.linalg.norm(b)
    assert rhoA + rhoB + rhoA * rhoB <= p["bound"]
    # exact integer product must be covered by the moduli
    ia_obj = [[int(v) for v in row.real] for row in ia]
    ib_obj = [[int(v) for v in row.real] for row in ib]
    biggest = max(abs(sum(x * y for x, y in zip(ra, rb))) for ra in ia_obj for rb in ib_obj)
    assert 2 * biggest < math.prod(ozaki.MODULI[:p["moduli"]])


@pytest.mark.parametrize("comps", [1, 2])
def test_min_moduli_is_a_lower_bound(planner, comps):
    """The automatic mode skips the statistics pass when even min_moduli(tol)
    moduli cannot beat FP64: no plan may ever need fewer. Random statistics,
    and the ones that make the row scale factor smallest (every component at
    the row maximum, maxima just below a power of two) or the rows sparsest
    (one non-zero component, a power of two)."""
    rng = np.random.default_rng(40 + comps)
    seen = 0
    for trial in range(400):
        k = int(rng.integers(1, 400000))
        tol = float(10.0 ** rng.uniform(-15.5, -6))
        rows = int(rng.integers(1, 64))
        kind = trial % 3
        if kind == 0:
            mA, sA = random_stats(rng, rows, k, comps, int(rng.integers(0, 40)))
            mB, sB = random_stats(rng, rows + 3, k, comps, int(rng.integers(0, 40)))
        elif kind == 1:  # dense rows: scale factor -> 1/2
            e = rng.integers(-20, 20, size=rows)
            mA = np.nextafter(np.exp2(e + 1.0), 0.0)
            sA = mA ** 2 * k * comps
            mB, sB = mA[::-1].copy(), sA[::-1].copy()
        else:  # one non-zero component per row
            mA = np.exp2(rng.integers(-20, 20, size=rows).astype(np.float64))
            sA = mA ** 2
            mB, sB = mA.copy(), sA.copy()
        p = run_plan(planner, mA, sA, mB, sB, k, comps, tol)
        if p["reason"] in (0, 4):  # planned, or only the moduli cap refused it
            seen += 1
            assert p["moduli"] >= planner.min_moduli_c(tol), (k, tol, kind)
    assert seen > 300
    # the bound itself: 2 ceil(log2(1.01 / tol)) bits of the moduli list
    for tol in (1e-13, 1e-14, 3e-9):
        bits = 2 * math.ceil(math.log2(1.01 / tol))
        need = next(n for n in range(1, 21) if sum(math.log2(m) for m in ozaki.MODULI[:n]) >= bits)
        assert planner.min_moduli_c(tol) == need


def test_plan_fallbacks(planner):
    ones = np.ones(4)
    ok = run_plan(planner, ones, ones * 100, ones, ones * 100, 100, 1, 1e-14)
    assert ok["ok"] and ok["reason"] == 0
    nan = run_plan(planner, np.array([1.0, np.nan, 1, 1]), ones, ones, ones, 100, 1, 1e-14)
    assert not nan["ok"] and nan["reason"] == 1
    inf = run_plan(planner, ones, np.array([1, np.inf, 1, 1.0]), ones, ones, 100, 1, 1e-14)
    assert not inf["ok"] and inf["reason"] == 1
    zero = run_plan(planner, np.zeros(4), np.zeros(4), ones, ones, 100, 1, 1e-14)
    assert not zero["ok"] and zero["reason"] == 2
    # each row holds one entry of magnitude 1 and k - 1 negligible ones: the
    # rounding bound needs P = log2(sqrt(k) / tol) > 53 bits
    wide = run_plan(planner, ones, ones, ones, ones, 10 ** 6, 1, 1e-14)
    assert not wide["ok"] and wide["reason"] == 3
    few = run_plan(planner, ones, ones * 100, ones, ones * 100, 100, 1, 1e-14, max_moduli=3)
    assert not few["ok"] and few["reason"] == 4


@pytest.mark.parametrize("cplx", [False, True])
def test_emulate_pairs_matches_emulate(cplx):
    """The sampled-entry form of the restatement equals the full one everywhere."""
    rng = np.random.default_rng(21 + cplx)
    m, n, k = 9, 6, 40
    A = rng.uniform(-1, 1, k * m) * np.exp2(rng.integers(-20, 20, k * m))
    B = rng.uniform(-1, 1, k * n)
    if cplx:
        A = A + 1j * rng.uniform(-1, 1, k * m)
        B = B + 1j * rng.uniform(-1, 1, k * n)
    opA = 2 if cplx else 1
    C = ozaki.emulate(opA, 0, m, n, k, A, k, B, k, 37, 44)
    opa = ozaki.op_rows(A, k, opA, m, k)
    opbT = B.reshape(n, k)
    pairs = [(i, j) for i in range(m) for j in range(n)]
    got = ozaki.emulate_pairs([opa[i] for i, _ in pairs], [opbT[j] for _, j in pairs], 37, 44)
    assert all(got[s] == C[i, j] for s, (i, j) in enumerate(pairs))


def i8_shape(lib, m, n, k, G, cus=256):
    kpad = (k + 255) // 256 * 256
    out = np.zeros(6, dtype=np.int64)
    sec = lib.i8_shape_c(m, n, kpad, G, cus, out.ctypes.data)
    return dict(cfg=int(out[0]), bm=int(out[1]), bn=int(out[2]), tiles=int(out[3]),
                splits=int(out[4]), kSplit=int(out[5]), kpad=kpad, seconds=sec)


@pytest.mark.parametrize("m,n,k,G,cfg,splits", [
    (2048, 2048, 200000, 42, 0, 2),   # headline: 256 x 256 tiles, 21 full waves
    (2048, 2048, 200000, 14, 0, 2),   # real product, 14 moduli: 7 full waves
    (2048, 2048, 200000, 15, 0, 4),   # real product, 15 moduli: 4 splits fill the last wave
    (100000, 4096, 4096, 39, 0, 1),   # sbs100k: one reduction split, uint8 output
    (512, 512, 200000, 42, 0, 6),     # mid-size: more splits of 256 x 256 tiles
    (1024, 512, 100000, 42, 0, 3),    # 336 tiles: 3 splits fill 4 waves (2 would leave 1.5 of 3)
    (100, 90, 50000, 42, 1, 16),      # small: 128 x 128 tiles (256-row padding would be 4x),
                                      # 13 splits reach 2 x 256 blocks, 16 fill two waves
])
def test_int8_shape_choice(planner, m, n, k, G, cfg, splits):
    s = i8_shape(planner, m, n, k, G)
    assert (s["cfg"], s["splits"]) == (cfg, splits), s
    assert s["bm"] == s["bn"] == (256 if cfg == 0 else 128)
    assert s["tiles"] == -(-m // s["bm"]) * -(-n // s["bn"])


@pytest.mark.parametrize("seed", range(4))
def test_int8_shape_invariants(planner, seed):
    """Every shape the planner returns covers the reduction with whole 256-byte
    splits within the INT32 range, and is the cheaper of the two modelled shapes."""
    rng = np.random.default_rng(seed)
    for _ in range(200):
        m, n = (int(x) for x in rng.integers(1, 5000, size=2))
        k = int(rng.integers(1, 2_000_000))
        G = int(rng.choice([1, 3])) * int(rng.integers(8, 21))
        s = i8_shape(planner, m, n, k, G)
        assert s["kSplit"] % 256 == 0 and 0 < s["kSplit"] <= (1 << 17) - 256
        assert s["splits"] * s["kSplit"] >= s["kpad"] > (s["splits"] - 1) * s["kSplit"]
        assert s["seconds"] > 0
