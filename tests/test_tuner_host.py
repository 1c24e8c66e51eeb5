"""Host tests of the FP64 kernels' tile planner (spla_amd/csrc/runtime/tuner.cpp).

* every plan covers the reduction: splits x kSplit >= the embedded k, with no
  empty split, kSplit a whole number of 16-element k-blocks;
* a forced split count (SPLA_AMD_SPLITS) is honoured exactly, clamped to the
  planner's maximum, for every count from 1 to 2000 (a count outside the planner's
  own candidate sequence once left a plan with kSplit = 0);
* whole-device plans of HBM-bound shapes (small m, n, long k) split far enough
  to occupy every CU, while the plans for the block products of a batch keep at
  most 64 splits;
* the headline shape keeps its compute-bound plan.
"""

import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@pytest.fixture(scope="module")
def tuner(tmp_path_factory):
    out = tmp_path_factory.mktemp("tuner") / "libtuner.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC",
                    "-I", str(ROOT / "spla_amd" / "csrc"), "-I", str(ROOT / "include"),
                    str(ROOT / "spla_amd" / "csrc" / "runtime" / "tuner.cpp"),
                    str(ROOT / "tests" / "cpp" / "tuner_capi.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    ll = ctypes.c_longlong
    lib.plan_gemm_c.argtypes = [ctypes.c_int, ll, ll, ll, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.POINTER(ll)]
    return lib


def plan(lib, cplx, m, n, k, force_cfg=-1, force_splits=0, whole=True, cus=256):
    out = (ctypes.c_longlong * 5)()
    lib.plan_gemm_c(int(cplx), m, n, k, cus, force_cfg, force_splits, int(whole), out)
    return dict(cfg=out[0], splits=out[1], kSplit=out[2], tilesM=out[3], tilesN=out[4])


def covers(p, cplx, k):
    kp = (2 if cplx else 1) * k
    return (p["splits"] >= 1 and p["kSplit"] > 0 and p["kSplit"] % 16 == 0
            and p["splits"] * p["kSplit"] >= kp and (p["splits"] - 1) * p["kSplit"] < kp)


@pytest.mark.parametrize("cplx", [False, True])
def test_plans_cover_the_reduction(tuner, cplx):
    for m in (1, 16, 33, 300, 2048, 100000):
        for n in (1, 16, 64, 700, 4096):
            for k in (1, 17, 4096, 200000, 8000000):
                for whole in (True, False):
                    p = plan(tuner, cplx, m, n, k, whole=whole)
                    assert covers(p, cplx, k), (m, n, k, whole, p)


def test_forced_split_count_is_honoured(tuner):
    m, n, k = 16, 16, 8_000_000
    limit = plan(tuner, True, m, n, k)["splits"]
    seen = set()
    for s in range(1, 2001):
        p = plan(tuner, True, m, n, k, force_splits=s)
        assert covers(p, True, k), (s, p)
        seen.add(p["splits"])
        assert p["splits"] <= max(s, limit)
    assert max(seen) >= 1000 and 1024 in seen  # counts outside the candidate sequence too


def test_hbm_bound_shapes_fill_the_device(tuner):
    p = plan(tuner, True, 16, 16, 8_000_000)
    assert p["tilesM"] * p["tilesN"] * p["splits"] >= 256  # every CU streams a share
    q = plan(tuner, True, 16, 16, 8_000_000, whole=False)
    assert q["splits"] <= 64  # batch blocks share the device
    h = plan(tuner, True, 2048, 2048, 200000)
    assert h["splits"] == 1 and h["tilesM"] * h["tilesN"] == 512  # 128 x 128 embedded tiles


@pytest.mark.parametrize("cplx", [False, True])
def test_gemm_grid_reaches_every_kblock_count_of_small_tiles(tuner, cplx):
    """The GEMM reference grid (m, n, k in 1, 13, 32, 263; every operation pair) runs the
    64 x 64 tile, whose loads go two k-blocks ahead, with splits of one and of two
    k-blocks (the loop's prologue edges) and with odd and even longer loops."""
    sizes = (1, 13, 32, 263)
    counts = set()
    for m in sizes:
        for n in sizes:
            for k in sizes:
                p = plan(tuner, cplx, m, n, k)
                if p["cfg"] != 1:  # TileM
                    continue
                kp = -(-(2 if cplx else 1) * k // 16) * 16
                for s in range(p["splits"]):
                    left = kp - s * p["kSplit"]
                    if left > 0:
                        counts.add(min(p["kSplit"], left) // 16)
    assert {1, 2} <= counts
    assert any(c >= 3 and c % 2 for c in counts) and any(c >= 4 and c % 2 == 0 for c in counts)


@pytest.mark.parametrize("cplx", [False, True])
def test_single_precision_short_reductions_reach_small_tile_prologue(tuner, cplx):
    """The short-reduction shapes of the single-precision bit-exact test (m = 13,
    n = 32, k = 8, 16, 24) run the 64 x 64 tile with splits of one and of two k-blocks."""
    counts = set()
    for k in (8, 16, 24):
        p = plan(tuner, cplx, 13, 32, k)
        assert p["cfg"] == 1, (k, p)  # TileM
        kp = -(-(2 if cplx else 1) * k // 16) * 16
        counts.add(min(p["kSplit"], kp) // 16)
    assert {1, 2} <= counts
