"""Host-side parts of bench.py (no GPU):

* --gpus above 1 is refused with exit status 2 and a message (the engine runs on
  one GPU as a single rank);
* the bench's own accuracy check (error-free products and compensated pairwise
  sums) agrees with the double-double oracle on the same dot products, and with
  exact rational arithmetic on small integer-scaled inputs, so the sampled
  normwise error it reports is itself accurate.
"""

from __future__ import annotations

import subprocess
import sys
from fractions import Fraction
from pathlib import Path

import numpy as np

import oracle

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr + out.stdout


def test_accurate_dot_matches_double_double_oracle():
    rng = np.random.default_rng(5)
    k, rows = 3000, 6
    x = rng.uniform(-1, 1, (rows, k)) * np.exp2(rng.integers(-30, 30, (rows, k)))
    y = rng.uniform(-1, 1, (rows, k))
    got = bench._accurate_dot(x, y)
    for r in range(rows):
        ref = np.zeros(1)
        # oracle: C (1x1) = A^T B with A, B k x 1, double-double accumulation
        oracle.gemm(1, 0, 1, 1, k, 1.0, np.ascontiguousarray(x[r]), k,
                    np.ascontiguousarray(y[r]), k, 0.0, ref, 1, accurate=True)
        assert abs(got[r] - ref[0]) <= 2 * np.spacing(abs(ref[0])), (r, got[r], ref[0])


def test_accurate_dot_exact_on_cancelling_sums():
    # values whose exact dot product cancels to a tiny result: naive summation fails
    rng = np.random.default_rng(9)
    k = 400
    x = rng.integers(-2**20, 2**20, k).astype(np.float64) * 2.0**30
    y = rng.integers(-2**20, 2**20, k).astype(np.float64)
    x = np.concatenate([x, -x, [3.0]])
    y = np.concatenate([y, y, [0.5]])
    exact = sum(Fraction(a) * Fraction(b) for a, b in zip(x, y))
    got = bench._accurate_dot(x[None, :], y[None, :])[0]
    assert got == float(exact) == 1.5


def test_accurate_products_grid_matches_double_double_oracle():
    # the bench's accuracy gate runs accurate_products on the GPU; the same torch
    # code on CPU tensors against the oracle (complex, conjugated rows of A^H)
    import torch

    rng = np.random.default_rng(11)
    k, R, C = 2500, 5, 4
    A = (rng.uniform(-1, 1, (R, k)) + 1j * rng.uniform(-1, 1, (R, k))) * np.exp2(
        rng.integers(-20, 20, (R, k)))
    B = rng.uniform(-1, 1, (C, k)) + 1j * rng.uniform(-1, 1, (C, k))
    got = bench.accurate_products(torch.from_numpy(A).conj().resolve_conj(),
                                  torch.from_numpy(B), rows_per_pass=2).numpy()
    for r in range(R):
        for c in range(C):
            ref = np.zeros(1, dtype=np.complex128)
            oracle.gemm(2, 0, 1, 1, k, 1.0, np.ascontiguousarray(A[r]), k,
                        np.ascontiguousarray(B[c]), k, 0.0, ref, 1, accurate=True)
            assert abs(got[r, c].real - ref[0].real) <= 2 * np.spacing(abs(ref[0].real))
            assert abs(got[r, c].imag - ref[0].imag) <= 2 * np.spacing(abs(ref[0].imag))


def test_accurate_products_exact_on_cancelling_sums():
    import torch

    rng = np.random.default_rng(9)
    k = 300
    x = rng.integers(-2**20, 2**20, k).astype(np.float64) * 2.0**30
    y = rng.integers(-2**20, 2**20, k).astype(np.float64)
    X = torch.from_numpy(np.concatenate([x, -x, [3.0]])[None, :])
    Y = torch.from_numpy(np.concatenate([y, y, [0.5]])[None, :])
    assert bench.accurate_products(X, Y)[0, 0].item() == 1.5


def test_sampled_grid_rows_and_columns_are_distinct():
    rows, cols = bench.sampled_rows_cols(2048, 2048, 64, 3)
    assert len(set(rows)) == 64 == len(set(cols))
    rows, cols = bench.sampled_rows_cols(13, 2048, 64, 3)
    assert sorted(rows) == list(range(13))
