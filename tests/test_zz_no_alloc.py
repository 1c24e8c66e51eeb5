"""spla_amd_zgemm_noalloc through the C ABI on the GPU (no counterpart in SPLA:
the allocation-free variant of spla_zgemm for callers that cannot allocate).

Checked against spla_zgemm on the same data in FP64 mode, which makes the same
plan on the same kernels: the results must be bit-identical. Two calls in a
row, the second with a smaller product (the shared scratch array keeps the
first call's partial sums), then a product whose planned partial sums exceed
the 64 MiB array (SPLA_GPU_ALLOCATION_ERROR, C untouched) and host operands
(SPLA_GPU_INVALID_DEVICE_POINTER_ERROR). The context's device-memory pool
stays empty: nothing is allocated.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
import spla_amd as spla
from spla_amd import _lib
from helpers import make_buffer, rel_error, to_device

pytestmark = pytest.mark.gpu


def zgemm(fn, ctx, opA, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    al = np.array([alpha], dtype=np.complex128)
    be = np.array([beta], dtype=np.complex128)
    ptr = lambda x: x if x is None or isinstance(x, int) else (
        x.data_ptr() if hasattr(x, "data_ptr") else x.ctypes.data)
    return fn(opA, 0, m, n, k, al.ctypes.data, ptr(A), lda, ptr(B), ldb, be.ctypes.data, ptr(C),
              ldc, ctx.handle)


def test_noalloc_matches_zgemm(device):
    import torch

    lib = _lib.lib()
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_gemm_mode(spla.GemmMode.FP64)
    cases = [(300, 260, 50000, 2, 0.5 - 1j, 0.25 + 0.5j),   # 34 reduction splits: 40 MiB
             (40, 24, 20000, 1, 1.0, 0.0)]                   # smaller, 209 splits: 3 MiB
    for i, (m, n, k, opA, alpha, beta) in enumerate(cases):
        print(f"noalloc case {i}: m={m} n={n} k={k}", flush=True)
        lda, ldc = k + 3, m + 2
        A = make_buffer(41, 1 + i, k, m, lda, np.complex128)
        B = make_buffer(41, 5 + i, k, n, k, np.complex128)
        C0 = make_buffer(41, 9 + i, m, n, ldc, np.complex128)
        dA, dB = to_device(A, device), to_device(B, device)
        c_ref, c_new = to_device(C0, device), to_device(C0, device)
        torch.cuda.synchronize()
        assert zgemm(lib.spla_zgemm, ctx, opA, m, n, k, alpha, dA, lda, dB, k, beta, c_ref,
                     ldc) == 0
        before = ctx.allocated_memory_gpu()
        assert zgemm(lib.spla_amd_zgemm_noalloc, ctx, opA, m, n, k, alpha, dA, lda, dB, k, beta,
                     c_new, ldc) == 0
        assert ctx.allocated_memory_gpu() == before
        got, ref = c_new.cpu().numpy(), c_ref.cpu().numpy()
        assert np.array_equal(got, ref), i
        exact = C0.copy()
        oracle.gemm(opA, 0, m, n, k, alpha, A, lda, B, k, beta, exact, ldc)
        assert rel_error(got, exact, A[:lda * m], B) <= 1e-13


def test_noalloc_limits(device):
    import torch

    lib = _lib.lib()
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    # m = n = 800, k = 50000: 28 reduction splits, 273 MiB of partial sums
    m = n = 800
    k = 50000
    print("noalloc limit case: m=n=800 k=50000", flush=True)
    dA = torch.zeros(k * m, dtype=torch.complex128, device=device)
    dB = torch.zeros(k * n, dtype=torch.complex128, device=device)
    dC = torch.full((m * n,), 3.0 + 1j, dtype=torch.complex128, device=device)
    torch.cuda.synchronize()
    err = zgemm(lib.spla_amd_zgemm_noalloc, ctx, 2, m, n, k, 1.0, dA, k, dB, k, 0.0, dC, m)
    assert err == spla.SplaError.GPU_ALLOCATION_ERROR
    assert bool((dC == 3.0 + 1j).all())
    # host operands would need staging buffers
    hA = np.zeros(64 * 8, dtype=np.complex128)
    dC8 = torch.zeros(64, dtype=torch.complex128, device=device)
    dB8 = torch.zeros(64 * 8, dtype=torch.complex128, device=device)
    err = zgemm(lib.spla_amd_zgemm_noalloc, ctx, 2, 8, 8, 64, 1.0, hA, 64, dB8, 64, 0.0, dC8, 8)
    assert err == spla.SplaError.GPU_INVALID_DEVICE_POINTER_ERROR
    # m == 0 returns at once; invalid parameters as spla_zgemm
    assert zgemm(lib.spla_amd_zgemm_noalloc, ctx, 2, 0, 8, 64, 1.0, None, 0, None, 0, 0.0, None,
                 0) == 0
    assert zgemm(lib.spla_amd_zgemm_noalloc, ctx, 2, 8, 8, 64, 1.0, dB8, 63, dB8, 64, 0.0, dC8,
                 8) == spla.SplaError.INVALID_PARAMETER_ERROR
    assert ctx.allocated_memory_gpu() == 0
