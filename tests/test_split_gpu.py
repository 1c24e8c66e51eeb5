"""Split-precision (Ozaki scheme II, INT8 matrix cores) products on the GPU.

* bit-exact against the exact integer restatement oracle/ozaki.py (alpha = 1,
  beta = 0), for every operand layout, real and complex: this pins the
  residues, the INT8 GEMM, the CRT reconstruction and the conversion;
* normwise error ||C - C_dd|| / (||A|| ||B||) against the double-double oracle
  is below the plan's reported bound and below 1e-13, including the
  exponent-spread inputs uniform[-1,1] * 2^e, e in [-10, 10];
* split-K (k > 2^17 - 128 forces INT32 split sums), fill modes, beta != 0,
  host-resident operands, run-to-run determinism;
* fall-back to FP64 for non-finite inputs, zero operands, too wide exponent
  ranges and, in automatic mode, when FP64 is expected to be faster.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import spla_amd as spla
from helpers import OPS, make_buffer, op_shape, rel_error, to_device, to_matrix
from oracle import ozaki

pytestmark = pytest.mark.gpu


@pytest.fixture()
def sctx(device):
    c = spla.Context(spla.ProcessingUnit.GPU)
    c.set_gemm_mode(spla.GemmMode.SPLIT_INT8)
    return c


def spread_buffer(seed, mid, rows, cols, ld, dtype, spread=10):
    rng = np.random.default_rng(seed * 17 + mid)
    e = rng.integers(-spread, spread + 1, size=ld * cols)
    v = rng.uniform(-1, 1, size=ld * cols) * np.exp2(e)
    if np.dtype(dtype).kind == "c":
        v = v + 1j * rng.uniform(-1, 1, size=ld * cols) * np.exp2(e)
    return v.astype(dtype)


def gpu_gemm(ctx, device, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    dC = to_device(C, device)
    spla.gemm(opA, opB, m, n, k, alpha, to_device(A, device), lda, to_device(B, device), ldb, beta,
              dC, ldc, ctx)
    return dC.cpu().numpy()


LAYOUTS = [("C", "N"), ("T", "N"), ("N", "N"), ("C", "C"), ("N", "T"), ("T", "C")]


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("opA,opB", LAYOUTS)
@pytest.mark.parametrize("shape", [(13, 11, 70), (37, 29, 300), (70, 3, 9)])
def test_split_bit_exact_against_integer_restatement(sctx, device, dtype, opA, opB, shape):
    if dtype == np.float64 and "C" in (opA, opB):
        pytest.skip("conjugation is the identity for real operands")
    m, n, k = shape
    oa, ob = OPS[opA], OPS[opB]
    ra, ca = op_shape(oa, m, k)
    rb, cb = op_shape(ob, k, n)
    A = make_buffer(3, 1, ra, ca, ra + 2, dtype)
    B = make_buffer(3, 2, rb, cb, rb + 1, dtype)
    C0 = np.zeros(m * n, dtype=dtype)
    got = gpu_gemm(sctx, device, oa, ob, m, n, k, 1.0, A, ra + 2, B, rb + 1, 0.0, C0, m)
    info = sctx.last_product_info()
    assert info["split"], info
    exp = ozaki.emulate(oa, ob, m, n, k, A, ra + 2, B, rb + 1, info["bits_a"], info["bits_b"])
    assert np.array_equal(got.reshape(n, m).T, exp)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
@pytest.mark.parametrize("opA,opB", LAYOUTS)
@pytest.mark.parametrize("data", ["uniform", "spread"])
def test_split_accuracy(sctx, device, dtype, opA, opB, data):
    if dtype == np.float64 and "C" in (opA, opB):
        pytest.skip("conjugation is the identity for real operands")
    m, n, k = 150, 130, 4000
    oa, ob = OPS[opA], OPS[opB]
    ra, ca = op_shape(oa, m, k)
    rb, cb = op_shape(ob, k, n)
    mk = spread_buffer if data == "spread" else make_buffer
    A = mk(5, 1, ra, ca, ra, dtype)
    B = mk(5, 2, rb, cb, rb, dtype)
    C0 = make_buffer(5, 3, m, n, m, dtype)
    alpha = (0.75 - 0.5j) if dtype == np.complex128 else 0.75
    beta = (0.5 + 0.25j) if dtype == np.complex128 else 0.5
    got = gpu_gemm(sctx, device, oa, ob, m, n, k, alpha, A, ra, B, rb, beta, C0, m)
    info = sctx.last_product_info()
    assert info["split"], info
    ref = C0.copy()
    oracle.gemm(oa, ob, m, n, k, alpha, A, ra, B, rb, beta, ref, m, accurate=True)
    err = rel_error(got, ref, A, B) / abs(alpha)
    assert err <= info["bound"] + 1e-16, (err, info)
    assert err <= 1e-13


@pytest.mark.parametrize("case", ["z_out8", "z_int32", "d_out8"])
def test_split_large_tiles_bit_exact(sctx, device, case):
    """256 x 256 INT8 tiles (the configuration of large products), uint8 residue
    output (one reduction split) and INT32 split sums (k > 2^17 - 256): sampled
    entries equal the exact integer restatement bit for bit."""
    import torch

    cplx = case.startswith("z")
    m = n = 1024 if cplx else 1792  # tiles x products >= 2 x 256: large tiles
    k = 140000 if case == "z_int32" else (1500 if cplx else 700)
    gen = torch.Generator(device=device)
    gen.manual_seed(4321)

    def operand(rows):
        x = torch.rand(rows * k * (2 if cplx else 1), generator=gen, device=device,
                       dtype=torch.float64) * 2 - 1
        return torch.view_as_complex(x.view(-1, 2)) if cplx else x

    dA, dB = operand(m), operand(n)
    dC = torch.zeros(m * n, dtype=dA.dtype, device=device)
    opA = 2 if cplx else 1
    spla.gemm(opA, 0, m, n, k, 1.0, dA, k, dB, k, 0.0, dC, m, sctx)
    info = sctx.last_product_info()
    assert info["split"], info
    rng = np.random.default_rng(17)
    rows, cols = rng.integers(0, m, 32), rng.integers(0, n, 32)
    a = dA.view(m, k)[torch.from_numpy(rows).to(device)].cpu().numpy()
    b = dB.view(n, k)[torch.from_numpy(cols).to(device)].cpu().numpy()
    if cplx:
        a = np.conj(a)
    want = ozaki.emulate_pairs(list(a), list(b), info["bits_a"], info["bits_b"])
    got = dC.view(n, m).cpu().numpy()[cols, rows]
    assert all(g == w for g, w in zip(got, want)), case


@pytest.mark.parametrize("case", ["z_large_out8", "d_large_out8", "z_small_int32", "d_small_out8"])
def test_int8_gemm_mfma_forms_bit_identical(sctx, device, monkeypatch, case):
    """The INT8 GEMM's two matrix-core forms (16 x 16 x 64, the default, and
    32 x 32 x 32, SPLA_AMD_I8_MFMA=32) give the same product bit for bit over the
    whole output: 256 x 256 tiles with uint8 residue output, and 128 x 128 tiles
    with uint8 output and with INT32 split sums (k > 2^17 - 256)."""
    import torch

    cplx = case.startswith("z")
    m, n, k = {"z_large_out8": (1024, 1024, 1500), "d_large_out8": (1792, 1536, 700),
               "z_small_int32": (72, 40, 140000), "d_small_out8": (130, 70, 3000)}[case]
    gen = torch.Generator(device=device)
    gen.manual_seed(97)
    dtype = torch.complex128 if cplx else torch.float64

    def operand(rows):
        x = torch.rand(rows * k * (2 if cplx else 1), generator=gen, device=device,
                       dtype=torch.float64) * 2 - 1
        return torch.view_as_complex(x.view(-1, 2)) if cplx else x

    dA, dB = operand(m), operand(n)
    out = {}
    for form in ("16", "32"):
        monkeypatch.setenv("SPLA_AMD_I8_MFMA", form)
        dC = torch.zeros(m * n, dtype=dtype, device=device)
        spla.gemm(2 if cplx else 1, 0, m, n, k, 1.0, dA, k, dB, k, 0.0, dC, m, sctx)
        assert sctx.last_product_info()["split"]
        out[form] = dC.cpu().numpy()
    assert np.array_equal(out["16"], out["32"]), case


@pytest.mark.parametrize("dtype", [np.float64, np.complex128], ids=["d", "z"])
def test_split_k_beyond_int32_range(sctx, device, dtype):
    m, n, k = 40, 24, 140000  # > 2^17 - 128: INT32 sums are split
    A = make_buffer(8, 1, k, m, k, dtype)
    B = make_buffer(8, 2, k, n, k, dtype)
    C0 = np.zeros(m * n, dtype=dtype)
    got = gpu_gemm(sctx, device, 2 if dtype == np.complex128 else 1, 0, m, n, k, 1.0, A, k, B, k,
                   0.0, C0, m)
    info = sctx.last_product_info()
    assert info["split"]
    ref = C0.copy()
    oracle.gemm(2 if dtype == np.complex128 else 1, 0, m, n, k, 1.0, A, k, B, k, 0.0, ref, m,
                accurate=True)
    assert rel_error(got, ref, A, B) <= min(info["bound"] + 1e-16, 1e-13)


def test_split_deterministic(sctx, device):
    m, n, k = 200, 100, 20000
    A = to_device(make_buffer(9, 1, k, m, k, np.complex128), device)
    B = to_device(make_buffer(9, 2, k, n, k, np.complex128), device)
    outs = []
    for _ in range(2):
        dC = to_device(np.zeros(m * n, dtype=np.complex128), device)
        spla.gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.0, dC, m, sctx)
        outs.append(dC.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("fill", [spla.FillMode.UPPER, spla.FillMode.LOWER])
def test_split_ssbtr_fill(sctx, device, fill):
    m, n, k, ro, co = 600, 520, 900, 7, 300
    ldc = m + ro + 5
    A = make_buffer(12, 1, k, m, k, np.complex128)
    B = make_buffer(12, 2, k, n, k, np.complex128)
    C0 = make_buffer(12, 3, ldc, n + co, ldc, np.complex128)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 64, 64)
    dC = to_device(C0, device)
    spla.pgemm_ssbtr(m, n, k, 2, 1.0, to_device(A, device), k, to_device(B, device), k, 0.0, dC,
                     ldc, ro, co, fill, dist, sctx)
    assert sctx.last_product_info()["split"]
    got = to_matrix(dC.cpu().numpy(), ldc, n + co, ldc)
    ref_full = C0.copy()
    oracle.gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.0, ref_full[ro + co * ldc:], ldc, accurate=True)
    R = to_matrix(ref_full, ldc, n + co, ldc)
    rows = np.arange(ldc)[:, None]
    cols = np.arange(n + co)[None, :]
    inside = (rows >= ro) & (rows < ro + m) & (cols >= co)
    req = inside & ((cols >= rows) if fill == spla.FillMode.UPPER else (rows >= cols))
    assert rel_error(got[req], R[req], A, B) <= 1e-13
    O = to_matrix(C0, ldc, n + co, ldc)
    assert np.array_equal(got[~inside], O[~inside])


def test_split_host_resident_operands(sctx, device):
    m, n, k = 128, 96, 6000
    A = make_buffer(13, 1, k, m, k, np.complex128)
    B = make_buffer(13, 2, k, n, k, np.complex128)
    C = make_buffer(13, 3, m, n, m, np.complex128)
    ref = C.copy()
    sctx.set_stream_chunk_bytes(1 << 20)  # several k-chunks
    dist = spla.MatrixDistribution.create_mirror(spla.MPI_COMM_SELF)
    spla.pgemm_ssb(m, n, k, 2, 1.0, A, k, B, k, 0.5, C, m, 0, 0, dist, sctx)
    oracle.gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.5, ref, m, accurate=True)
    assert rel_error(C, ref, A, B) <= 1e-13


def test_split_fallbacks(sctx, device):
    m, n, k = 20, 16, 50000
    A = make_buffer(14, 1, k, m, k, np.float64)
    B = make_buffer(14, 2, k, n, k, np.float64)
    C0 = np.zeros(m * n)
    # non-finite input: FP64 kernels (NaN propagates as in FP64)
    An = A.copy()
    An[5] = np.nan
    got = gpu_gemm(sctx, device, 1, 0, m, n, k, 1.0, An, k, B, k, 0.0, C0, m)
    info = sctx.last_product_info()
    assert not info["split"] and info["reason"] == 1
    G = got.reshape(n, m)
    assert np.isnan(G[:, 0]).all() and np.isfinite(G[:, 1:]).all()
    # zero operand
    gpu_gemm(sctx, device, 1, 0, m, n, k, 1.0, np.zeros_like(A), k, B, k, 0.0, C0, m)
    info = sctx.last_product_info()
    assert not info["split"] and info["reason"] == 2
    # one dominant entry per row: the normwise bound needs more than 53 bits
    Aw = np.full(k * m, 1e-40)
    Aw[::k] = 1.0
    got = gpu_gemm(sctx, device, 1, 0, m, n, k, 1.0, Aw, k, B, k, 0.0, C0, m)
    info = sctx.last_product_info()
    assert not info["split"] and info["reason"] == 3
    # k = 1, real: the error target k * 2^-53 needs more than 53 bits per entry
    got1 = gpu_gemm(sctx, device, 1, 0, m, n, 1, 1.0, A, k, B, k, 0.0, C0, m)
    info = sctx.last_product_info()
    assert not info["split"] and info["reason"] == 3
    exp1 = np.outer(A[::k][:m], B[::k][:n]).T.ravel()
    assert np.array_equal(got1, exp1)
    ref = C0.copy()
    oracle.gemm(1, 0, m, n, k, 1.0, Aw, k, B, k, 0.0, ref, m, accurate=True)
    assert rel_error(got, ref, Aw, B) <= 1e-13


def test_auto_mode_keeps_small_products_on_fp64(device):
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_gemm_mode(spla.GemmMode.AUTO)
    m, n, k = 32, 32, 64
    A = make_buffer(15, 1, k, m, k, np.complex128)
    B = make_buffer(15, 2, k, n, k, np.complex128)
    gpu_gemm(ctx, device, 2, 0, m, n, k, 1.0, A, k, B, k, 0.0, np.zeros(m * n, complex), m)
    info = ctx.last_product_info()
    assert not info["split"] and info["reason"] == 6
    ctx.set_gemm_mode(spla.GemmMode.FP64)
    gpu_gemm(ctx, device, 2, 0, m, n, k, 1.0, A, k, B, k, 0.0, np.zeros(m * n, complex), m)
    assert ctx.last_product_info()["reason"] == 7


def test_phase_timing_reports_split_phases(sctx, device):
    sctx.set_timing(True)
    m, n, k = 256, 256, 8192
    A = to_device(make_buffer(16, 1, k, m, k, np.complex128), device)
    B = to_device(make_buffer(16, 2, k, n, k, np.complex128), device)
    dC = to_device(np.zeros(m * n, dtype=np.complex128), device)
    spla.gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.0, dC, m, sctx)
    t = sctx.last_timing()
    assert t["split_int8_gemm"] > 0 and t["split_residues"] > 0 and t["split_crt"] > 0
    assert t["gemm_fp64"] == 0
    sctx.set_timing(False)


def test_split_workspace_limit_chunks_the_reduction(device):
    # A workspace limit below the product's planes: the reduction runs in chunks,
    # each a split product of its own added into C in FP64 (beta on the first).
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_gemm_mode(spla.GemmMode.SPLIT_INT8)
    m, n, k = 128, 120, 20000
    A = make_buffer(17, 1, k, m, k, np.complex128)
    B = make_buffer(17, 2, k, n, k, np.complex128)
    C0 = make_buffer(17, 3, m, n, m, np.complex128)
    ref = C0.copy()
    oracle.gemm(2, 0, m, n, k, 0.5 + 1j, A, k, B, k, 0.25, ref, m, accurate=True)
    dA, dB = to_device(A, device), to_device(B, device)
    outs = []
    for limit in (0, 256 << 20, 16 << 20):
        ctx.set_split_workspace_limit(limit)
        dC = to_device(C0, device)
        spla.gemm(2, 0, m, n, k, 0.5 + 1j, dA, k, dB, k, 0.25, dC, m, ctx)
        info = ctx.last_product_info()
        got = dC.cpu().numpy()
        assert rel_error(got, ref, A, B) <= 1e-13, (limit, info)
        outs.append((got, info))
    (whole, iw), (chunked, ic), (fp64, i8) = outs
    assert iw["split"] and ic["split"]
    # two chunks of 10240 reduction indices: the rounding differs from the whole product
    assert not np.array_equal(whole, chunked)
    # 16 MiB holds fewer than 4096 reduction indices: FP64 kernels, reason 8
    assert not i8["split"] and i8["reason"] == 8


def test_split_streamed_chunks_share_one_workspace(device):
    # Host-resident operands in split precision: every streamed k-chunk is a split
    # product of its own; they reuse one device workspace, so the device memory
    # the context holds does not grow with the number of chunks.
    held = []
    for k in (8000, 32000):
        ctx = spla.Context(spla.ProcessingUnit.GPU)
        ctx.set_gemm_mode(spla.GemmMode.SPLIT_INT8)
        ctx.set_stream_chunk_bytes(4 << 20)
        m, n = 160, 96
        A = make_buffer(18, 1, k, m, k, np.complex128)
        B = make_buffer(18, 2, k, n, k, np.complex128)
        C = make_buffer(18, 3, m, n, m, np.complex128)
        ref = C.copy()
        spla.gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.5, C, m, ctx)
        assert ctx.last_product_info()["split"]
        oracle.gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.5, ref, m, accurate=True)
        assert rel_error(C, ref, A, B) <= 1e-13
        held.append(ctx.allocated_memory_gpu())
    assert held[1] <= 1.25 * held[0], held
