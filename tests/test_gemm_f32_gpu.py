"""Single precision (spla_sgemm / spla_cgemm and the ps / pc distributed forms) on the
FP32 matrix cores, FP32 accumulation as in the vendor s/c gemm SPLA calls.

* integer-valued inputs (|entries| <= 8, every partial sum an integer below 2^24, so
  FP32 arithmetic is exact): bit-exact against the exact product for every operand
  combination, the small and the large tile configurations, split-K shapes, padded
  leading dimensions, host and device operands;
* uniform inputs: normwise error ||C - C_dd|| / (||A|| ||B||) against the
  double-double oracle of the same (single-precision) data <= 1e-6 (FP32 rounding of
  the products, the sums and the result; the FP64 kernels would give ~1e-16, a wrong
  layout ~1);
* pgemm_ssbtr in complex single precision on a block-cyclic C (lower fill).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import spla_amd as spla
from helpers import OPS, make_buffer, op_shape, rel_error, to_device, to_matrix

pytestmark = pytest.mark.gpu

WIDE = {np.dtype(np.float32): np.float64, np.dtype(np.complex64): np.complex128}


def case(ctx, device, dtype, opA, opB, m, n, k, kind, where="device", alpha=2.0, beta=3.0,
         seed=17):
    oa, ob = OPS[opA], OPS[opB]
    ra, ca = op_shape(oa, m, k)
    rb, cb = op_shape(ob, k, n)
    lda, ldb, ldc = ra + 3, rb + 1, m + 2
    A = make_buffer(seed, 1, ra, ca, lda, dtype, kind)
    B = make_buffer(seed, 2, rb, cb, ldb, dtype, kind)
    C0 = make_buffer(seed, 3, m, n, ldc, dtype, kind)
    wide = WIDE[np.dtype(dtype)]
    ref = C0.astype(wide)
    oracle.gemm(oa, ob, m, n, k, alpha, A.astype(wide), lda, B.astype(wide), ldb, beta, ref, ldc,
                accurate=True)
    if where == "device":
        dC = to_device(C0, device)
        spla.gemm(oa, ob, m, n, k, alpha, to_device(A, device), lda, to_device(B, device), ldb,
                  beta, dC, ldc, ctx)
        got = dC.cpu().numpy()
    else:
        got = C0.copy()
        spla.gemm(oa, ob, m, n, k, alpha, A, lda, B, ldb, beta, got, ldc, ctx)
    return A, B, got, ref, (ra, ca, lda), (rb, cb, ldb), ldc


COMBOS = [("N", "N"), ("C", "N"), ("T", "C"), ("N", "T"), ("C", "C")]
# (13, 32, 8 / 16 / 24): the 64 x 64 tile (loads two k-blocks ahead) with one and two
# k-blocks in a split, real and complex
SHAPES = [(13, 32, 263), (263, 5, 40), (300, 200, 150), (1024, 768, 96), (40, 24, 20000),
          (13, 32, 8), (13, 32, 16), (13, 32, 24)]


@pytest.mark.parametrize("dtype", [np.float32, np.complex64], ids=["s", "c"])
@pytest.mark.parametrize("opA,opB", COMBOS)
def test_f32_integer_inputs_bit_exact(ctx, device, dtype, opA, opB):
    if dtype == np.float32 and "C" in (opA, opB):
        pytest.skip("conjugation is the identity for real operands")
    for m, n, k in SHAPES:
        _, _, got, ref, _, _, ldc = case(ctx, device, dtype, opA, opB, m, n, k, "int")
        g, r = to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc)
        assert np.array_equal(g.astype(r.dtype), r), (opA, opB, m, n, k)
        # the ld padding of C is untouched
        pad = got[: ldc * n].reshape(n, ldc)[:, m:]
        assert np.all(pad == 1234.5)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64], ids=["s", "c"])
@pytest.mark.parametrize("opA,opB", [("C", "N"), ("N", "N"), ("T", "T"), ("N", "T")])
def test_f32_large_tile_every_kblock_count_bit_exact(ctx, device, dtype, opA, opB,
                                                     monkeypatch):
    """The 128 x 128 tile (split-barrier schedule in every layout but real MN-major B),
    forced, on 1 to 5 k-blocks of 16 embedded reduction elements in one and in two
    splits, edge tiles in m and n; integer inputs, bit-exact."""
    if dtype == np.float32:
        opA = "T" if opA == "C" else opA
    monkeypatch.setenv("SPLA_AMD_TILE", "L")
    per_kblock = 8 if dtype == np.complex64 else 16
    for splits in ("1", "2"):
        monkeypatch.setenv("SPLA_AMD_SPLITS", splits)
        for kb in (1, 2, 3, 4, 5):
            for m, n in ((130, 140), (64, 128)):
                k = kb * per_kblock
                _, _, got, ref, _, _, ldc = case(ctx, device, dtype, opA, opB, m, n, k, "int")
                g, r = to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc)
                assert np.array_equal(g.astype(r.dtype), r), (splits, kb, m, n)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64], ids=["s", "c"])
@pytest.mark.parametrize("where", ["device", "host"])
def test_f32_accuracy(ctx, device, dtype, where):
    for opA, opB in (("C", "N"), ("N", "N"), ("T", "C")):
        for m, n, k in ((150, 130, 20000), (1024, 512, 300)):
            A, B, got, ref, sa, sb, ldc = case(ctx, device, dtype, opA, opB, m, n, k, "uniform",
                                               where)
            err = rel_error(to_matrix(got, m, n, ldc), to_matrix(ref, m, n, ldc),
                            to_matrix(A, *sa), to_matrix(B, *sb))
            assert err <= 1e-6, (opA, opB, m, n, k, where, err)


def test_pcgemm_ssbtr_block_cyclic(ctx, device):
    import torch

    m, n, k, rb, cb = 300, 260, 900, 64, 48
    A = make_buffer(23, 1, k, m, k, np.complex64, "int")
    B = make_buffer(23, 2, k, n, k, np.complex64, "int")
    C0 = make_buffer(23, 3, m, n, m, np.complex64, "int")
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, rb, cb)
    dC = to_device(C0, device)
    spla.pgemm_ssbtr(m, n, k, spla.Operation.CONJ_TRANSPOSE, 1.0, to_device(A, device), k,
                     to_device(B, device), k, 2.0, dC, m, 0, 0, spla.FillMode.LOWER, dist, ctx)
    torch.cuda.synchronize()
    got = to_matrix(dC.cpu().numpy(), m, n, m).astype(np.complex128)
    ref = C0.astype(np.complex128).copy()
    oracle.gemm(2, 0, m, n, k, 1.0, A.astype(np.complex128), k, B.astype(np.complex128), k, 2.0,
                ref, m)
    ref = to_matrix(ref, m, n, m)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    need = ii >= jj  # every element of the lower triangle is computed (bit-exact: integers)
    assert np.array_equal(got[need], ref[need])
