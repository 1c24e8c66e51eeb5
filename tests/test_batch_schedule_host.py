"""Host checks of the batched launch's schedule (spla_amd/csrc/runtime/batch_schedule.cpp)
and of the kernels' indexing over it (spla_amd/csrc/kernels/batched.hip), before
any GPU run:

* the GEMM launch holds, for every tile of every block that SPLA's fill predicate
  (src/block_generation/block.hpp:58-66) marks active, the splits 0 .. splits - 1
  once each, and nothing for inactive tiles; the splits cover the embedded
  reduction [0, Kp) exactly, in whole 16-element k-blocks, none empty;
* every operand, C and partial-sum address the GEMM and the reduction kernels
  form, replayed here from the descriptors as the kernels form them, lies inside
  its buffer (A with lda = k (ssb) or m (sbs), B with ldb = k, C with ldc = m,
  the workspace of workDoubles doubles); the reduction launch holds exactly the
  active tiles of the blocks with more than one split, once each;
* both grids fit a launch; blocks with m, n or k = 0, and alpha = 0 (no
  reduction: C <- beta C), are handled;
* the persistent launch's static lists (one per resident workgroup slot)
  partition the items, each in dispatch order;
* on the 512-block ring-step workload the split counts and tile choice are
  those the cost model prefers, and the model time is finite.
"""

from __future__ import annotations

import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")

TILE = {0: (128, 128), 1: (64, 64)}  # TileL, TileM: embedded rows x columns
FULL, UPPER, LOWER = 0, 1, 2


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("bs") / "libbs.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC",
                    "-I", str(ROOT / "spla_amd" / "csrc"), "-I", str(ROOT / "include"),
                    str(ROOT / "spla_amd" / "csrc" / "runtime" / "batch_schedule.cpp"),
                    str(ROOT / "tests" / "cpp" / "batch_schedule_capi.cpp"), "-o", str(out)],
                   check=True)
    h = ctypes.CDLL(str(out))
    ll = ctypes.c_longlong
    h.bs_plan.restype = ctypes.c_void_p
    h.bs_plan.argtypes = [ctypes.POINTER(ll), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    h.bs_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ll)]
    h.bs_model_seconds.restype = ctypes.c_double
    h.bs_model_seconds.argtypes = [ctypes.c_void_p]
    h.bs_blocks.argtypes = [ctypes.c_void_p, ctypes.POINTER(ll)]
    h.bs_items.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    h.bs_reduce.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    h.bs_num_slots.restype = ctypes.c_longlong
    h.bs_num_slots.argtypes = [ctypes.c_void_p]
    h.bs_slots.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                           ctypes.POINTER(ctypes.c_int)]
    h.bs_free.argtypes = [ctypes.c_void_p]
    return h


def plan(lib, shapes, cplx, cfg, scale=1.0, alpha_zero=False, cus=256, eff=0.75):
    arr = np.ascontiguousarray(np.array(shapes, dtype=np.int64).reshape(-1, 6))
    h = lib.bs_plan(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(shapes), int(cplx),
                    int(alpha_zero), cfg, scale, cus, eff)
    cnt = (ctypes.c_longlong * 4)()
    lib.bs_counts(h, cnt)
    blocks = np.zeros((len(shapes), 7), dtype=np.int64)
    lib.bs_blocks(h, blocks.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    items = np.zeros((cnt[0], 4), dtype=np.int32)
    lib.bs_items(h, items.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    red = np.zeros((cnt[1], 3), dtype=np.int32)
    lib.bs_reduce(h, red.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    model = lib.bs_model_seconds(h)
    nslots = lib.bs_num_slots(h)
    start = np.zeros(nslots + 1, dtype=np.int32)
    order = np.zeros(max(1, cnt[0]), dtype=np.int32)
    lib.bs_slots(h, start.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                 order.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    lib.bs_free(h)
    return dict(blocks=blocks, items=items, reduce=red, work=cnt[2], cfg=cnt[3], model=model,
                slot_start=start, slot_items=order[:cnt[0]], cus=cus)


def active(fill, ro, co, r, c, nr, nc):
    if fill == UPPER:
        return c + nc + co > r + ro
    if fill == LOWER:
        return r + nr + ro > c + co
    return True


def emulate(p, shapes, cplx, sbs, alpha_zero=False):
    """Replays the kernels' index arithmetic over the schedule (batched.hip)."""
    comps = 2 if cplx else 1
    bm, bn = TILE[p["cfg"]]
    ar = bm // comps
    blocks, items, red, work = p["blocks"], p["items"], p["reduce"], p["work"]
    assert len(items) < 2**31 and len(red) < 2**31
    seen = {}
    for b, tm, tn, s in items:
        assert 0 <= b < len(shapes)
        m, n, k = shapes[b][:3]
        bm_, bn_, Kp, Mp, splits, kSplit, wo = blocks[b]
        assert (bm_, bn_) == (m, n) and m > 0 and n > 0
        assert Kp == (0 if alpha_zero else k * comps) and Mp == m * comps
        assert 0 <= tm < -(-m // ar) and 0 <= tn < -(-n // bn) and 0 <= s < splits
        assert kSplit > 0 and kSplit % 16 == 0
        k_begin = s * kSplit
        k_end = min(Kp, k_begin + kSplit)
        assert k_begin < k_end or (Kp == 0 and s == 0)
        rows = min(ar, m - tm * ar)
        cols = min(bn, n - tn * bn)
        assert rows > 0 and cols > 0
        if k_end > k_begin:
            # A: ssb k x m (K-major, lda = k): element (l, i) at i * lda + l;
            # sbs m x k (MN-major, lda = m): element (i, l) at l * lda + i
            i_max, l_max = tm * ar + rows - 1, (k_end - 1) // comps
            a_last = (i_max + l_max * m) if sbs else (i_max * k + l_max)
            assert a_last < m * k
            assert (tn * bn + cols - 1) * k + l_max < n * k  # B k x n, ldb = k
        if splits > 1:
            last = wo + s * n * Mp + (tn * bn + cols - 1) * Mp + comps * (tm * ar + rows) - 1
            assert wo >= 0 and last < work
        else:
            assert (tn * bn + cols - 1) * m + tm * ar + rows - 1 < m * n  # C, ldc = m
        seen.setdefault((b, tm, tn), []).append((s, k_begin, k_end))
    # tiles: exactly the active ones, every split once, the splits tile [0, Kp)
    expect = set()
    for b, (m, n, k, fill, ro, co) in enumerate(shapes):
        if m <= 0 or n <= 0:
            assert not any(items[:, 0] == b) if len(items) else True
            continue
        for tm in range(-(-m // ar)):
            for tn in range(-(-n // bn)):
                r, c = tm * ar, tn * bn
                if active(fill, ro, co, r, c, min(ar, m - r), min(bn, n - c)):
                    expect.add((b, tm, tn))
    assert set(seen) == expect
    for (b, tm, tn), parts in seen.items():
        splits, Kp = blocks[b][4], blocks[b][2]
        parts.sort()
        assert [q[0] for q in parts] == list(range(splits))
        edges = [q[1] for q in parts] + [parts[-1][2]]
        assert edges[0] == 0 and edges[-1] == Kp
        assert all(parts[i][2] == parts[i + 1][1] for i in range(len(parts) - 1))
    # persistent launch: the slots' static lists partition the items, each list in
    # dispatch order, one slot per resident workgroup at most
    start, order = p["slot_start"], p["slot_items"]
    per_cu = 4 if p["cfg"] == 1 else 2
    assert len(start) - 1 == max(1, min(p["cus"] * per_cu, len(items)))
    assert start[0] == 0 and start[-1] == len(items) and (np.diff(start) >= 0).all()
    assert sorted(order.tolist()) == list(range(len(items)))
    for w in range(len(start) - 1):
        mine = order[start[w]:start[w + 1]]
        assert (np.diff(mine) > 0).all()
    red_set = [tuple(t) for t in red]
    assert len(red_set) == len(set(red_set))
    assert set(red_set) == {t for t in expect if blocks[t[0]][4] > 1}
    # workspaces of different blocks do not overlap
    spans = sorted((blocks[b][6], blocks[b][6] + blocks[b][4] * blocks[b][1] * blocks[b][3])
                   for b in range(len(shapes)) if blocks[b][4] > 1)
    assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1))
    assert not spans or spans[-1][1] <= work


def ragged(seed, count, fills=True):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m, n, k = int(rng.integers(1, 300)), int(rng.integers(1, 220)), int(rng.integers(0, 2000))
        fill = int(rng.integers(0, 3)) if fills else FULL
        out.append((m, n, k, fill, int(rng.integers(0, 400)), int(rng.integers(0, 400))))
    out += [(0, 17, 30, FULL, 0, 0), (19, 0, 30, FULL, 0, 0), (23, 31, 0, FULL, 0, 0),
            (1, 1, 1, FULL, 0, 0), (256, 256, 4096, LOWER, 0, 0), (512, 300, 60000, UPPER, 7, 3)]
    return out


@pytest.mark.parametrize("cfg", [0, 1])
@pytest.mark.parametrize("cplx", [True, False])
@pytest.mark.parametrize("sbs", [False, True])
def test_kernel_indexing_over_ragged_blocks(lib, cfg, cplx, sbs):
    shapes = ragged(41, 40, fills=not sbs)
    for scale in (0.5, 1.0, 2.0):
        p = plan(lib, shapes, cplx, cfg, scale)
        assert p["cfg"] == cfg
        emulate(p, shapes, cplx, sbs)


def test_alpha_zero_reads_no_operand(lib):
    shapes = ragged(5, 20)
    p = plan(lib, shapes, True, 0, alpha_zero=True)
    emulate(p, shapes, True, False, alpha_zero=True)
    assert (p["blocks"][:, 4] <= 1).all() and len(p["reduce"]) == 0


def batch512_shapes():
    rng = np.random.default_rng(5)
    cnt = 512
    ms = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), cnt))).astype(int)
    ns = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), cnt))).astype(int)
    ks = np.rint(np.exp(rng.uniform(np.log(4096), np.log(65536), cnt))).astype(int)
    return [(int(ms[i]), int(ns[i]), int(ks[i]), LOWER if i % 2 else FULL, 0, 0)
            for i in range(cnt)]


@pytest.mark.parametrize("cfg", [0, 1])
def test_ring_step_workload_schedule(lib, cfg):
    shapes = batch512_shapes()
    p = plan(lib, shapes, True, cfg)
    emulate(p, shapes, True, False)
    assert 0 < p["model"] < 1.0
    # long reductions are split, short ones are not split below 16 k-blocks each
    kp = p["blocks"][:, 2]
    assert (p["blocks"][:, 4] <= np.maximum(1, (kp // 16) // 16)).all()
    assert p["blocks"][:, 4].max() > 1
    # the items leave each slot about three workgroups' worth of work or more
    assert len(p["items"]) >= 3 * 256 * (2 if cfg == 0 else 4) * 0.5


def test_splitting_few_large_blocks_fills_the_device(lib):
    few = [(512, 512, 60000, FULL, 0, 0)] * 2  # 64 tiles of 128 x 128 for 512 slots
    whole = plan(lib, few, True, 0, 1e-6)     # one split per tile
    split = plan(lib, few, True, 0, 1.0)
    assert whole["blocks"][0][4] == 1 and len(whole["items"]) == 64
    assert len(split["items"]) >= 3 * 512 and len(split["reduce"]) == 64
    assert split["model"] < whole["model"] / 4
