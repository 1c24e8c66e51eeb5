"""Batched ring-step block products (spla_amd_batch_*) on the GPU.

The batch runs many independent C_b <- alpha op(A_b) B_b + beta C_b, the block
products of one pgemm_ssb / pgemm_ssbtr ring step
(src/pgemm_ssb/ring_ssb_gpu.cpp:193-197) or of one pgemm_sbs ring step
(op(A) = A, src/pgemm_sbs/ring_sbs_gpu.cpp:57-88). In FP64 mode a batch is one
GEMM launch over the tiles and reduction splits of every block plus one
fixed-order reduction launch, both replayed from a hipGraph; in split precision
the blocks run one by one on the single-product engine. Checked here:

* every block against the oracle (normwise error <= 1e-13), real and complex,
  opA transpose, conjugate transpose and (sbs) none, ragged sizes including m,
  n or k equal to zero, ld padding and beta != 0;
* fill modes with tile granularity (SPLA's block_is_active,
  src/block_generation/block.hpp:58-66): the requested triangle is exact, no
  element outside the block is written, and elements of the block outside the
  triangle hold either their old value or the product;
* blocks with long reductions (split partial sums reduced in a fixed order):
  results are bitwise identical over repeated graph replays, with unrelated GPU
  work in between, and across measured and modelled tuning;
* beta accumulation over replays, alpha = 0 (C <- beta C without reading A, B);
* the split-precision path of a batch (gemm mode SPLIT_INT8).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import spla_amd as spla
from helpers import make_buffer, rel_error, to_device, to_matrix

pytestmark = pytest.mark.gpu


def make_blocks(seed, sizes, dtype, device, fills=None, pad=(0, 0, 0)):
    """Device operands for blocks (m, n, k); returns (specs, host copies)."""
    import torch

    specs, host = [], []
    for i, (m, n, k) in enumerate(sizes):
        lda, ldb, ldc = max(1, k) + pad[0], max(1, k) + pad[1], max(1, m) + pad[2]
        A = make_buffer(seed, 3 * i + 1, k, m, lda, dtype)
        B = make_buffer(seed, 3 * i + 2, k, n, ldb, dtype)
        C = make_buffer(seed, 3 * i + 3, m, n, ldc, dtype)
        dA, dB, dC = (to_device(x, device) for x in (A, B, C))
        fill, ro, co = fills[i] if fills else (spla.FillMode.FULL, 0, 0)
        specs.append(dict(m=m, n=n, k=k, A=dA, lda=lda, B=dB, ldb=ldb, C=dC, ldc=ldc, fill=fill,
                          row_offset=ro, col_offset=co))
        host.append((A, B, C))
    torch.cuda.synchronize()
    return specs, host


def expected(spec, host, opA, alpha, beta):
    A, B, C = host
    ref = C.copy()
    m, n, k = spec["m"], spec["n"], spec["k"]
    if m and n:
        # plain FP64 reference: its own normwise error is ~eps / sqrt(k), far below the bar
        oracle.gemm(opA, 0, m, n, k, alpha, A, spec["lda"], B, spec["ldb"], beta, ref, spec["ldc"])
    return ref


def check_block(spec, host, opA, alpha, beta, tol=1e-13):
    A, B, C = host
    got = spec["C"].cpu().numpy()
    ref = expected(spec, host, opA, alpha, beta)
    m, n, ldc = spec["m"], spec["n"], spec["ldc"]
    if m == 0 or n == 0:
        assert np.array_equal(got, C)
        return
    # ld padding is never written
    G = got[: ldc * n].reshape(n, ldc).T
    O = C[: ldc * n].reshape(n, ldc).T
    assert np.array_equal(G[m:], O[m:])
    g, r, o = to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc), to_matrix(C, m, n, ldc)
    rows = np.arange(m)[:, None] + spec["row_offset"]
    cols = np.arange(n)[None, :] + spec["col_offset"]
    if spec["fill"] == spla.FillMode.UPPER:
        req = cols >= rows
    elif spec["fill"] == spla.FillMode.LOWER:
        req = rows >= cols
    else:
        req = np.ones((m, n), dtype=bool)
    scale = max(abs(alpha), 1e-300)
    assert rel_error(g[req], r[req], A, B) / scale <= tol
    # outside the triangle: the old value or the product (whole tiles are computed)
    rest = ~req
    if rest.any():
        old = g[rest] == o[rest]
        close = np.abs(g[rest] - r[rest]) <= 1e-12 * (np.abs(r[rest]) + 1.0)
        assert (old | close).all()


def ragged_sizes(rng, count, max_m, max_n, max_k):
    out = []
    for _ in range(count):
        out.append((int(rng.integers(1, max_m)), int(rng.integers(1, max_n)),
                    int(rng.integers(0, max_k))))
    return out


@pytest.mark.parametrize("persistent", [False, True], ids=["grouped", "persistent"])
@pytest.mark.parametrize("dtype,opA", [(np.complex128, 2), (np.complex128, 1), (np.float64, 1)],
                         ids=["zC", "zT", "dT"])
def test_batch_ragged_blocks(ctx, device, dtype, opA, persistent):
    rng = np.random.default_rng(41)
    sizes = ragged_sizes(rng, 30, 300, 220, 2000)
    sizes += [(0, 17, 30), (19, 0, 30), (23, 31, 0), (1, 1, 1), (256, 256, 4096)]
    cplx = dtype == np.complex128
    alpha = (0.75 - 0.5j) if cplx else 0.75
    beta = (0.5 + 0.25j) if cplx else -0.5
    specs, host = make_blocks(7, sizes, dtype, device, pad=(3, 1, 2))
    batch = spla.Batch(ctx, specs, alpha, beta, complex_=cplx, opA=opA, persistent=persistent)
    batch.execute()
    for s, h in zip(specs, host):
        check_block(s, h, opA, alpha, beta)
    info = batch.info()
    assert info["items"] >= 1 and info["workgroups"] >= 1 and info["model_ms"] > 0
    assert info["persistent"] == persistent
    batch.close()


def test_batch_beta_zero_ignores_nan_in_c(ctx, device):
    sizes = [(70, 50, 300), (130, 90, 1000)]
    specs, host = make_blocks(8, sizes, np.complex128, device)
    for s in specs:
        s["C"].fill_(complex(float("nan"), 0.0))
    batch = spla.Batch(ctx, specs, 1.0, 0.0)
    batch.execute()
    for s, (A, B, _) in zip(specs, host):
        m, n, k = s["m"], s["n"], s["k"]
        ref = np.zeros(s["ldc"] * n, dtype=np.complex128)
        oracle.gemm(2, 0, m, n, k, 1.0, A, s["lda"], B, s["ldb"], 0.0, ref, s["ldc"])
        got = to_matrix(s["C"].cpu().numpy(), m, n, s["ldc"])
        assert rel_error(got, to_matrix(ref, m, n, s["ldc"]), A, B) <= 1e-13


@pytest.mark.parametrize("fill", [spla.FillMode.UPPER, spla.FillMode.LOWER])
def test_batch_fill_modes(ctx, device, fill):
    # blocks of one triangular ring step: offsets place each block in C
    fills, sizes = [], []
    for bi in range(4):
        for bj in range(4):
            sizes.append((150, 130, 700))
            fills.append((fill, bi * 150, bj * 130))
    specs, host = make_blocks(9, sizes, np.complex128, device, fills=fills)
    batch = spla.Batch(ctx, specs, 1.0 + 0.5j, 0.5)
    batch.execute()
    for s, h in zip(specs, host):
        check_block(s, h, 2, 1.0 + 0.5j, 0.5)


@pytest.mark.parametrize("persistent", [False, True], ids=["grouped", "persistent"])
def test_batch_split_k_deterministic_over_replays(ctx, device, persistent):
    import torch

    sizes = [(256, 192, 60000), (64, 64, 90000), (300, 40, 5000), (96, 128, 30000)]
    specs, host = make_blocks(10, sizes, np.complex128, device)
    batch = spla.Batch(ctx, specs, 1.0, 0.0, persistent=persistent)
    info = batch.info()
    first = None
    for rep in range(4):
        # unrelated GPU work between replays
        x = torch.randn(1 << 20, device=device)
        torch.cuda.synchronize()
        batch.execute()
        out = [s["C"].cpu().numpy() for s in specs]
        assert float(x.sum()) == float(x.sum())
        if first is None:
            first = out
            for s, h in zip(specs, host):
                check_block(s, h, 2, 1.0, 0.0)
        else:
            for a, b in zip(first, out):
                assert np.array_equal(a, b), rep
    # split-K in use: more items than output tiles
    assert info["items"] > len(sizes)


def test_batch_beta_accumulates_over_replays(ctx, device):
    sizes = [(100, 80, 3000), (33, 65, 20000)]
    specs, host = make_blocks(11, sizes, np.float64, device)
    batch = spla.Batch(ctx, specs, 0.5, 1.0, complex_=False, opA=1)
    for _ in range(3):
        batch.execute()
    for s, (A, B, C) in zip(specs, host):
        ref = C.copy()
        for _ in range(3):
            oracle.gemm(1, 0, s["m"], s["n"], s["k"], 0.5, A, s["lda"], B, s["ldb"], 1.0, ref,
                        s["ldc"])
        m, n, ldc = s["m"], s["n"], s["ldc"]
        got = to_matrix(s["C"].cpu().numpy(), m, n, ldc)
        assert rel_error(got, to_matrix(ref, m, n, ldc), A, B) <= 3e-13


def test_batch_measured_tuning_matches_model(ctx, device):
    sizes = [(512, 512, 8000)] * 3 + [(128, 64, 20000)] * 5
    specs, host = make_blocks(12, sizes, np.complex128, device)
    specs2, _ = make_blocks(12, sizes, np.complex128, device)
    b1 = spla.Batch(ctx, specs, 1.0, 0.0, measure=True)
    b2 = spla.Batch(ctx, specs2, 1.0, 0.0)
    assert b1.info()["measured_ms"] > 0
    # tuning left C untouched
    for s, (_, _, C) in zip(specs, host):
        assert np.array_equal(s["C"].cpu().numpy(), C)
    b1.execute()
    b2.execute()
    for s, s2, h in zip(specs, specs2, host):
        check_block(s, h, 2, 1.0, 0.0)
        check_block(s2, h, 2, 1.0, 0.0)


def test_batch_split_precision(device):
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_gemm_mode(spla.GemmMode.SPLIT_INT8)
    sizes = [(200, 150, 3000), (64, 300, 9000)]
    specs, host = make_blocks(13, sizes, np.complex128, device)
    batch = spla.Batch(ctx, specs, 1.0 - 1.0j, 0.25)
    batch.execute()
    assert ctx.last_product_info()["split"]
    for s, h in zip(specs, host):
        check_block(s, h, 2, 1.0 - 1.0j, 0.25)


def test_batch_split_precision_sbs_blocks(device):
    """sbs blocks (A m x k, MN-major) in split precision: residues of an MN-major A."""
    import torch

    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_gemm_mode(spla.GemmMode.SPLIT_INT8)
    sizes = [(700, 96, 3000), (300, 200, 5000)]
    specs, host = [], []
    for i, (m, n, k) in enumerate(sizes):
        A = make_buffer(45, 3 * i + 1, m, k, m + 1, np.complex128)
        B = make_buffer(45, 3 * i + 2, k, n, k, np.complex128)
        C = make_buffer(45, 3 * i + 3, m, n, m, np.complex128)
        dA, dB, dC = (to_device(x, device) for x in (A, B, C))
        specs.append(dict(m=m, n=n, k=k, A=dA, lda=m + 1, B=dB, ldb=k, C=dC, ldc=m))
        host.append((A, B, C))
    torch.cuda.synchronize()
    batch = spla.Batch(ctx, specs, 0.5 + 0.5j, -1.0, opA=spla.Operation.NONE)
    batch.execute()
    assert ctx.last_product_info()["split"]
    for s, (A, B, C) in zip(specs, host):
        m, n, k = s["m"], s["n"], s["k"]
        ref = C.copy()
        oracle.gemm(0, 0, m, n, k, 0.5 + 0.5j, A, s["lda"], B, k, -1.0, ref, m, accurate=True)
        assert rel_error(s["C"].cpu().numpy(), ref, A, B) <= 1e-13
    batch.close()


def test_batch_rejects_host_pointers(ctx, device):
    A = np.zeros(64, dtype=np.complex128)
    specs, _ = make_blocks(14, [(8, 8, 8)], np.complex128, device)
    specs[0]["A"] = A
    with pytest.raises(spla.GPUInvalidDevicePointerError):
        spla.Batch(ctx, specs, 1.0, 0.0)


@pytest.mark.parametrize("persistent", [False, True], ids=["grouped", "persistent"])
@pytest.mark.parametrize("dtype", [np.complex128, np.float64], ids=["z", "d"])
def test_batch_sbs_blocks(ctx, device, dtype, persistent):
    """sbs blocks: C_b <- alpha A_b B_b + beta C_b with A_b m x k (lda >= m)."""
    import torch

    rng = np.random.default_rng(43)
    sizes = [(int(rng.integers(1, 400)), int(rng.integers(1, 200)), int(rng.integers(0, 3000)))
             for _ in range(12)] + [(300, 64, 30000), (5, 7, 0)]
    cplx = dtype == np.complex128
    alpha = (0.25 + 1j) if cplx else 1.5
    beta = (-0.5 + 0.5j) if cplx else 0.25
    specs, host = [], []
    for i, (m, n, k) in enumerate(sizes):
        lda, ldc = m + 2, m + 1
        A = make_buffer(44, 3 * i + 1, m, k, lda, dtype)
        B = make_buffer(44, 3 * i + 2, k, n, max(1, k), dtype)
        C = make_buffer(44, 3 * i + 3, m, n, ldc, dtype)
        dA, dB, dC = (to_device(x, device) for x in (A, B, C))
        specs.append(dict(m=m, n=n, k=k, A=dA, lda=lda, B=dB, ldb=max(1, k), C=dC, ldc=ldc))
        host.append((A, B, C))
    torch.cuda.synchronize()
    batch = spla.Batch(ctx, specs, alpha, beta, complex_=cplx, opA=spla.Operation.NONE,
                       persistent=persistent)
    batch.execute()
    for s, (A, B, C) in zip(specs, host):
        m, n, k, ldc = s["m"], s["n"], s["k"], s["ldc"]
        ref = C.copy()
        oracle.gemm(0, 0, m, n, k, alpha, A, s["lda"], B, s["ldb"], beta, ref, ldc)
        got = s["C"].cpu().numpy()
        pad = got[: ldc * n].reshape(n, ldc)[:, m:]
        assert np.array_equal(pad, C[: ldc * n].reshape(n, ldc)[:, m:])
        assert rel_error(to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc), A, B) <= 1e-13
    assert batch.info()["items"] > len(sizes)
    batch.close()


def test_batch_alpha_zero_scales_c_without_reading_operands(ctx, device):
    import torch

    sizes = [(70, 50, 300), (33, 200, 5000)]
    specs, host = make_blocks(15, sizes, np.complex128, device)
    for s in specs:
        s["A"].fill_(complex(float("nan"), 0.0))
        s["B"].fill_(complex(float("nan"), 0.0))
    torch.cuda.synchronize()
    batch = spla.Batch(ctx, specs, 0.0, 0.5 - 0.5j)
    batch.execute()
    for s, (_, _, C) in zip(specs, host):
        got = s["C"].cpu().numpy()
        assert np.array_equal(got, (0.5 - 0.5j) * C)
    batch.close()


def test_batch_automatic_mode_mixes_grouped_fp64_and_split(device):
    """Automatic mode: blocks whose shape rules the split form out run as one grouped
    FP64 launch, the others block by block (split precision where it is expected to be
    faster); every block meets the FP64 bar, and two executions agree bitwise."""
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_gemm_mode(spla.GemmMode.AUTO)
    sizes = [(40, 30, 5000), (64, 100, 20000), (200, 48, 9000), (1024, 1024, 40000),
             (96, 96, 3000), (3, 2, 100)]
    specs, host = make_blocks(16, sizes, np.complex128, device)
    batch = spla.Batch(ctx, specs, 1.0 - 0.5j, 0.0)
    batch.execute()
    # the 1024 x 1024 x 40000 block is the one whose split form pays off
    assert batch.info()["split_blocks"] == 1, batch.info()
    first = [s["C"].cpu().numpy() for s in specs]
    for s, h in zip(specs, host):
        check_block(s, h, 2, 1.0 - 0.5j, 0.0)
    batch.execute()
    for s, f in zip(specs, first):
        assert np.array_equal(s["C"].cpu().numpy(), f)
    batch.close()


def test_batch_automatic_mode_keeps_grouped_launch_when_split_does_not_pay(device):
    """Automatic mode on small blocks with long reductions: each 64 x 64 x 60000 block
    may beat its own stand-alone FP64 estimate split (few tiles, long reduction), but
    together their split lower bounds (1.2 ms) do not beat the grouped FP64 launch's
    estimate for them (0.2 ms), so every block stays in the grouped launch; the results
    then equal the FP64 mode's bit for bit."""
    sizes = [(64, 64, 60000)] * 4 + [(40, 30, 5000), (100, 80, 3000)]
    results = {}
    for mode in (spla.GemmMode.AUTO, spla.GemmMode.FP64):
        ctx = spla.Context(spla.ProcessingUnit.GPU)
        ctx.set_gemm_mode(mode)
        specs, host = make_blocks(23, sizes, np.complex128, device)
        batch = spla.Batch(ctx, specs, 1.0, 0.0)
        batch.execute()
        if mode == spla.GemmMode.AUTO:
            assert batch.info()["split_blocks"] == 0, batch.info()
        results[mode] = [s["C"].cpu().numpy() for s in specs]
        batch.close()
    for a, b in zip(results[spla.GemmMode.AUTO], results[spla.GemmMode.FP64]):
        assert np.array_equal(a, b)
