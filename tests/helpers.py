"""Shared helpers of the parity tests: column-major buffers with leading
dimensions, device round trips and oracle comparisons."""

from __future__ import annotations

import numpy as np

import oracle

OPS = {"N": 0, "T": 1, "C": 2}


def op_shape(op: int, rows: int, cols: int) -> tuple[int, int]:
    """Stored shape (rows, cols) of an operand read as op(X) with op(X) rows x cols."""
    return (rows, cols) if op == 0 else (cols, rows)


def make_buffer(seed: int, mid: int, rows: int, cols: int, ld: int, dtype, kind="uniform"):
    """Column-major buffer with leading dimension ld (padding filled with a sentinel)."""
    cplx = np.issubdtype(np.dtype(dtype), np.complexfloating)
    size = max(ld * cols, 1)
    if kind == "int":
        rng = np.random.default_rng(seed * 131 + mid)
        v = rng.integers(-8, 9, size=size).astype(np.float64)
        if cplx:
            v = v + 1j * rng.integers(-8, 9, size=size)
    else:
        v = oracle.uniform(seed, mid, size, complex_=cplx)
    buf = np.asarray(v, dtype=dtype).copy()
    if ld > rows:
        view = buf[: ld * cols].reshape(cols, ld)
        view[:, rows:] = 1234.5  # sentinel in the ld padding
    return buf


def to_matrix(buf, rows, cols, ld):
    return buf[: ld * cols].reshape(cols, ld).T[:rows, :]


def to_device(buf, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(buf)).to(device)


def rel_error(got, ref, A, B):
    nA = np.linalg.norm(np.asarray(A, dtype=np.complex128))
    nB = np.linalg.norm(np.asarray(B, dtype=np.complex128))
    return oracle.normwise_error(got, ref, nA, nB)
