"""Benchmark of the local GEMM engine on one MI355X.

Workload (one step): spla_pzgemm_ssb on one rank (MPI_COMM_SELF), complex
double, C <- alpha * A^H * B + beta * C with A (k = 200000 x m = 2048) and
B (k x n = 2048), C 2048 x 2048 block cyclic with 256 x 256 blocks on a 1 x 1
grid, alpha = 1, beta = 0.5; A, B and C hold values uniform in [-1, 1) per
component (splitmix64, seed 42, matrix ids 1/2/3) and are resident in HBM
when the timed region starts.

Forms: "plain" (FP64 MFMA kernels) and "split" (FP64-accurate split precision:
Ozaki scheme II on the INT8 matrix cores, falling back to FP64 when its error
bound cannot be met). The split form is planned for a guaranteed normwise error
bound of 1e-13 (--split-tolerance), the accuracy the metric requires; the
library's default tolerance is 1e-14 (one more modulus on this workload).

value: FP64-accurate TFLOP/s (8*m*n*k / seconds per step) of the fastest form
whose measured normwise error ||C - C_ref|| / (||A|| ||B||) is <= 1e-13, from a
64 x 64 grid of sampled entries (4096) against error-free products and
compensated sums computed on the device. Every form run is reported.

workloads: after the headline, the default run measures every other
configuration in the same process with fewer steps, one entry each: the headline
with A, B, C in pinned host memory (PCIe-inclusive; plain, split, automatic),
pzgemm_sbs with mLocal = 100000 (plain, split), the real A^T B with an exponent
spread (plain, split, and a forced FP64 fall-back), 512 uneven ring-step blocks
through the batch API (plain, split, automatic), the headline in complex single
precision and two HBM-bound small m, n products. --no-workloads skips them;
--workload NAME runs one of them alone.

roofline: the dominant kernel of that form (FP64 GEMM kernel, or the batched
INT8 GEMM of the split form), its duration from HIP events the library records
on its launch stream (per-phase timing of the context), against MI355X's
published peak (FP64 matrix 78.6 TFLOP/s; INT8 dense 5.0 POP/s).

cpu_baseline: SPLA's CPU path (one host-BLAS zgemm, as in
src/gemm/gemm_host.cpp:55-72) on a k-truncated sample of the same workload,
timed on this host's cores; "config1_full" times the CPU-runnable configuration
(m = n = 256, k = 4096) in full.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

M = N = 2048
K = 200_000
SEED = 42
ALPHA = 1.0 + 0.0j
BETA = 0.5 + 0.0j
FP64_PEAK_TFLOPS = 78.6
INT8_PEAK_TOPS = 5000.0
# HBM bytes (read + write) per launch of the dominant kernel of each form on this
# workload, from rocprofv3 FETCH_SIZE and WRITE_SIZE passes (bench.py
# --copy-calibration ...), each scaled by what the counter reads for a 1 GiB device
# copy (reads x 2.0, writes x 1.0). They change only when the kernels or the split
# plan (number of moduli) change.
HBM_TRAFFIC_BYTES = {"split": 51.66e9 + 1.48e9, "plain": 116.67e9 + 0.07e9}
PROFILED_MODULI = 14  # moduli of the split plan the counters were taken on
METRIC = "FP64-accurate TFLOP/s, zgemm A^H·B k=2e5 m=n=2048: split-prec vs plain FP64"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# synthetic inputs on the device: splitmix64(seed, matrix id, draw index)

_G = 0x9E3779B97F4A7C15
_M1 = 0xBF58476D1CE4E5B9
_M2 = 0x94D049BB133111EB


def _s64(x: int) -> int:
    x &= (1 << 64) - 1
    return x - (1 << 64) if x >= 1 << 63 else x


def uniform_device(seed: int, mid: int, count: int, complex_: bool, device, chunk=1 << 25):
    """Same values as oracle.uniform(seed, mid, count, complex_), generated on the device."""
    import torch

    draws = count * (2 if complex_ else 1)
    out = torch.empty(draws, dtype=torch.float64, device=device)
    base = _s64(seed * _G + (mid << 40) + _G)
    m1, m2 = _s64(_M1), _s64(_M2)

    def lsr(z, s):
        return (z >> s) & ((1 << (64 - s)) - 1)

    for c0 in range(0, draws, chunk):
        n = min(chunk, draws - c0)
        z = torch.arange(c0, c0 + n, dtype=torch.int64, device=device) + base
        z = (z ^ lsr(z, 30)) * m1
        z = (z ^ lsr(z, 27)) * m2
        z = z ^ lsr(z, 31)
        out[c0:c0 + n] = lsr(z, 11).to(torch.float64) * (2.0 ** -53) * 2.0 - 1.0
    return torch.view_as_complex(out.view(-1, 2)) if complex_ else out


# ---------------------------------------------------------------------------

def _two_prod(x, y):
    """Error-free product x * y = p + e (Dekker, Veltkamp splitting, no FMA needed).
    x, y: numpy arrays or torch tensors (CPU or GPU), float64."""

    def split(v):
        c = v * 134217729.0  # 2^27 + 1
        hi = c - (c - v)
        return hi, v - hi

    p = x * y
    xh, xl = split(x)
    yh, yl = split(y)
    e = ((xh * yh - p) + xh * yl + xl * yh) + xl * yl
    return p, e


def _accurate_sum(v):
    """Sum of v (last axis) with error-free pairwise TwoSum steps: the rounding
    errors of every addition are kept and added at the end. numpy or torch."""
    try:
        import torch
        is_torch = isinstance(v, torch.Tensor)
    except ImportError:  # pragma: no cover
        is_torch = False
    if not is_torch:
        import numpy as np
        v = np.asarray(v, dtype=np.float64)
    errs = None
    while v.shape[-1] > 1:
        if v.shape[-1] % 2:
            if is_torch:
                import torch
                v = torch.cat([v, torch.zeros(v.shape[:-1] + (1,), dtype=v.dtype,
                                              device=v.device)], dim=-1)
            else:
                v = np.concatenate([v, np.zeros(v.shape[:-1] + (1,))], axis=-1)
        a, b = v[..., 0::2], v[..., 1::2]
        s = a + b
        bb = s - a
        e = ((a - (s - bb)) + (b - bb)).sum(-1)
        errs = e if errs is None else errs + e
        v = s
    return v[..., 0] + (errs if errs is not None else 0.0)


def _accurate_dot(x, y):
    """sum_l x[..., l] * y[..., l] to about double-double accuracy (numpy or torch)."""
    p, e = _two_prod(x, y)
    try:
        import torch
        if isinstance(p, torch.Tensor):
            return _accurate_sum(torch.cat([p, e], dim=-1))
    except ImportError:  # pragma: no cover
        pass
    import numpy as np
    return _accurate_sum(np.concatenate([p, e], axis=-1))


def accurate_products(X, Y, rows_per_pass=None):
    """Every entry sum_l X[r, l] * Y[c, l] of X (R x K) and Y (C x K), float64 or
    complex128 torch tensors on one device, accurate to about double-double.
    Complex: Re = [Xr, -Xi] . [Yr, Yi], Im = [Xr, Xi] . [Yi, Yr] (real embedding)."""
    import torch

    if X.is_complex():
        Xr, Xi, Yr, Yi = X.real, X.imag, Y.real, Y.imag
        re = accurate_products(torch.cat([Xr, -Xi], 1), torch.cat([Yr, Yi], 1), rows_per_pass)
        im = accurate_products(torch.cat([Xr, Xi], 1), torch.cat([Yi, Yr], 1), rows_per_pass)
        return torch.complex(re, im)
    R, C, K = X.shape[0], Y.shape[0], X.shape[1]
    rb = rows_per_pass or max(1, min(R, (1 << 27) // max(1, C * K)))
    out = torch.empty(R, C, dtype=torch.float64, device=X.device)
    for r0 in range(0, R, rb):
        x = X[r0:r0 + rb, None, :]
        out[r0:r0 + rb] = _accurate_dot(x.expand(-1, C, -1), Y[None, :, :].expand(x.shape[0], -1, -1))
    return out


def sampled_rows_cols(m, n, count, seed):
    """count x count grid of sampled entries: sorted distinct rows and columns."""
    import numpy as np

    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(m, size=min(m, count), replace=False))
    cols = np.sort(rng.choice(n, size=min(n, count), replace=False))
    return rows, cols


def normwise_from_samples(got, ref, m, n, nAB):
    """||C - C_ref||_F / (||A|| ||B||) estimated from sampled entries (RMS x sqrt(m n))."""
    import torch

    d = (got - ref).abs().double() / nAB  # scaled first: squares of huge entries overflow
    return float(torch.sqrt((d * d).mean() * (m * n)).item())


def sampled_error(dA, dB, dC0, dC, k, m, n, samples=64, seed=3):
    """Normwise error estimate ||C - C_ref|| / (||A|| ||B||) of the headline product
    from a samples x samples grid of entries (4096 by default); C_ref = alpha A^H B +
    beta C0 with error-free products and compensated sums, on the device."""
    import torch

    rows, cols = sampled_rows_cols(m, n, samples, seed)
    ri = torch.from_numpy(rows).to(dA.device)
    ci = torch.from_numpy(cols).to(dA.device)
    X = dA.view(m, k)[ri].conj()  # rows of A^H
    Y = dB.view(n, k)[ci]
    ab = accurate_products(X, Y)  # (rows, cols)
    C0 = dC0.view(n, m)[ci][:, ri].T  # C0[i, j]
    ref = ALPHA * ab + BETA * C0
    got = dC.view(n, m)[ci][:, ri].T
    nAB = torch.linalg.vector_norm(dA).item() * torch.linalg.vector_norm(dB).item()
    return normwise_from_samples(got, ref, m, n, nAB), len(rows) * len(cols)


def cpu_baseline(m, n, k_full, target_seconds=12.0):
    """SPLA's CPU path (host BLAS zgemm A^H B) on a k-truncated sample."""
    import numpy as np

    import oracle

    info = oracle.cpu_path()
    threads = oracle.lib().cpu_path_get_threads()

    def run(kk):
        A = oracle.uniform(SEED, 1, kk * m, complex_=True)
        B = oracle.uniform(SEED, 2, kk * n, complex_=True)
        C = oracle.uniform(SEED, 3, m * n, complex_=True)
        oracle.cpu_path_gemm(2, 0, m, n, kk, ALPHA, A, kk, B, kk, BETA, C, m)  # warm-up
        t0 = time.perf_counter()
        oracle.cpu_path_gemm(2, 0, m, n, kk, ALPHA, A, kk, B, kk, BETA, C, m)
        return time.perf_counter() - t0

    k1 = 4096
    t1 = run(k1)
    k2 = int(min(k_full, max(k1, k1 * target_seconds / max(t1, 1e-3))) // 1024 * 1024)
    t2 = run(k2) if k2 > k1 else t1
    rate = 8.0 * m * n * k2 / t2 / 1e12
    one = ""
    if threads and threads > 1:
        oracle.lib().cpu_path_set_threads(1)
        try:
            t1s = run(k1)
        finally:
            oracle.lib().cpu_path_set_threads(threads)
        one = f"; 1 thread at k={k1}: {8.0 * m * n * k1 / t1s / 1e12:.4f} TFLOP/s"
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        cores = os.cpu_count()
    cfg1 = cpu_config1(threads)
    return {
        "value": round(rate, 4),
        "unit": "TFLOP/s",
        "cores": threads if threads and threads > 0 else cores,
        "kind": "port",
        "sample": (f"{info['vendor']} cblas_zgemm('C','N') m=n={m}, k={k2} (of {k_full}); "
                   f"{t2:.2f} s; extrapolated full-k step {t2 * k_full / k2:.1f} s; "
                   f"all threads at k={k1}: {8.0 * m * n * k1 / t1 / 1e12:.4f} TFLOP/s{one}; "
                   f"host cores visible {cores}"),
        "config1_full": cfg1,
    }


def cpu_config1(threads):
    """The CPU-runnable configuration in full: pzgemm_ssb on one rank, m = n = 256,
    k = 4096, A^H B, alpha = 1, beta = 0 (at one rank SPLA's CPU path is one host
    zgemm, src/gemm/gemm_host.cpp:55-72); 1 warm-up, median of 7, all threads and 1."""
    import numpy as np

    import oracle

    m = n = 256
    k = 4096
    A = oracle.uniform(SEED, 1, k * m, complex_=True)
    B = oracle.uniform(SEED, 2, k * n, complex_=True)

    def med(nthreads):
        if nthreads:
            oracle.lib().cpu_path_set_threads(nthreads)
        C = np.zeros(m * n, dtype=np.complex128)
        oracle.cpu_path_gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.0, C, m)
        ts = []
        for _ in range(7):
            t0 = time.perf_counter()
            oracle.cpu_path_gemm(2, 0, m, n, k, 1.0, A, k, B, k, 0.0, C, m)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    try:
        t_all = med(0)
        t_one = med(1)
    finally:
        if threads:
            oracle.lib().cpu_path_set_threads(threads)
    flops = 8.0 * m * n * k
    return {"workload": "spla_pzgemm_ssb m=n=256 k=4096 A^H B complex, alpha=1 beta=0, "
                        "64x64 blocks on 1x1, SPLA's CPU path (host zgemm)",
            "ms": round(t_all * 1e3, 3), "tflops": round(flops / t_all / 1e12, 4),
            "threads": threads, "ms_1_thread": round(t_one * 1e3, 3),
            "tflops_1_thread": round(flops / t_one / 1e12, 4)}


WORKLOADS = ("ssb2048", "sbs100k", "ssb_real_spread", "batch512", "ssb2048_host",
             "ssb2048_c64", "ssb_small_mn")
# side workloads of the default run, in order (each a key of the JSON line's "workloads")
DEFAULT_SIDE = ("ssb2048_host", "sbs100k", "ssb_real_spread", "batch512", "ssb2048_c64",
                "ssb_small_mn")
SIDE_FORMS = {"ssb2048_host": ["plain", "split", "auto"], "sbs100k": ["plain", "split"],
              "ssb_real_spread": ["plain", "split"], "batch512": ["plain", "split", "auto"],
              "ssb2048_c64": ["plain"], "ssb_small_mn": ["plain"]}


def _timed(fn, steps, warmup, stream):
    """Runs fn warmup + steps times; returns (wall ms per step, event ms per step)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps * 1e3
    return wall, e0.elapsed_time(e1) / steps


def _phases(ctx, fn) -> dict:
    """Per-phase device times (ms) of one extra call of fn."""
    import torch

    ctx.set_timing(True)
    fn()
    torch.cuda.synchronize()
    ph = ctx.last_timing()
    ctx.set_timing(False)
    return {k: round(v, 3) for k, v in ph.items() if v > 0}


def _modes(spla):
    return {"plain": spla.GemmMode.FP64, "split": spla.GemmMode.SPLIT_INT8,
            "auto": spla.GemmMode.AUTO}


def _info_keys(info) -> dict:
    return {"split_used": info["split"], "fallback_reason": info["reason"],
            "moduli": info["moduli"], "bits": [info["bits_a"], info["bits_b"]],
            "error_bound": info["bound"]}


def _release(*tensors) -> None:
    import torch

    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _ssb_reference(dA, dB, k, m, n, rows, cols, conj=True):
    """Accurate A^H B (conj) or A^T B entries at rows x cols (A k x m, B k x n)."""
    import torch

    ri = torch.from_numpy(rows).to(dA.device)
    ci = torch.from_numpy(cols).to(dA.device)
    X = dA.view(m, k)[ri]
    if conj and X.is_complex():
        X = X.conj()
    return accurate_products(X.resolve_conj(), dB.view(n, k)[ci])


def _gather(dC, m, n, rows, cols):
    """Entries C[rows, cols] of a column-major m x n device matrix, as (rows, cols)."""
    import torch

    ri = torch.from_numpy(rows).to(dC.device)
    ci = torch.from_numpy(cols).to(dC.device)
    return dC.view(n, m)[ci][:, ri].T


def workload_ssb2048_host(spla, dev, stream, steps, warmup, forms, tol):
    """the headline product with A, B and C in pinned host memory
    (k-chunked streaming over PCIe overlapped with the kernels). PCIe-inclusive."""
    import torch

    m, n, k = M, N, K
    hA = torch.empty(k * m, dtype=torch.complex128, pin_memory=True)
    hB = torch.empty(k * n, dtype=torch.complex128, pin_memory=True)
    hC = torch.empty(m * n, dtype=torch.complex128, pin_memory=True)
    dA = uniform_device(SEED, 1, k * m, True, dev)
    hA.copy_(dA)
    nA = torch.linalg.vector_norm(dA).item()
    rows, cols = sampled_rows_cols(m, n, 32, 7)
    ab = None
    dB = uniform_device(SEED, 2, k * n, True, dev)
    hB.copy_(dB)
    nB = torch.linalg.vector_norm(dB).item()
    ab = _ssb_reference(dA, dB, k, m, n, rows, cols)
    del dA, dB
    dC0 = uniform_device(SEED, 3, m * n, True, dev)
    hC0 = dC0.cpu()
    ref = ALPHA * ab + BETA * _gather(dC0, m, n, rows, cols)
    del dC0
    _release()
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_stream(stream)
    ctx.set_split_tolerance(tol)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 256,
                                                             256)
    flops = 8.0 * m * n * k
    out = {"config": {"workload": "spla_pzgemm_ssb A^H B complex m=n=2048 k=200000, beta=0.5, "
                                  "A, B, C in pinned host memory (PCIe-inclusive)",
                      "split_tolerance": tol},
           "dtype": "c128", "forms": {}}
    for form in forms:
        ctx.set_gemm_mode(_modes(spla)[form])
        run = lambda: spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, ALPHA, hA, k, hB, k,
                                     BETA, hC, m, 0, 0, dist, ctx)
        hC.copy_(hC0)
        run()
        got = _gather(hC.to(dev), m, n, rows, cols)
        err = normwise_from_samples(got, ref, m, n, nA * nB)
        info = ctx.last_product_info()
        wall, _ = _timed(run, steps, warmup, stream)
        out["forms"][form] = {"tflops_pcie_inclusive": round(flops / wall / 1e9, 3),
                              "ms_per_step": round(wall, 3), "normwise_error": err,
                              "error_samples": len(rows) * len(cols), **_info_keys(info)}
        log(f"ssb2048_host {form}: {out['forms'][form]}")
    ctx.close()
    del hA, hB, hC
    _release()
    return out


def workload_sbs100k(spla, dev, stream, steps, warmup, forms, tol):
    """pzgemm_sbs, mLocal = 100000, n = k = 4096, device-resident."""
    import torch

    mL, n, k = 100000, 4096, 4096
    dA = uniform_device(SEED, 1, mL * k, True, dev)
    dB = uniform_device(SEED, 2, k * n, True, dev)
    dC = torch.zeros(mL * n, dtype=torch.complex128, device=dev)
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_stream(stream)
    ctx.set_split_tolerance(tol)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 128,
                                                             128)
    flops = 8.0 * mL * n * k
    rows, cols = sampled_rows_cols(mL, n, 32, 3)
    X = dA.view(k, mL)[:, torch.from_numpy(rows).to(dev)].T.contiguous()  # rows of A
    Y = dB.view(n, k)[torch.from_numpy(cols).to(dev)]                     # columns of B
    ref = accurate_products(X, Y)
    nAB = torch.linalg.vector_norm(dA).item() * torch.linalg.vector_norm(dB).item()
    out = {"config": {"workload": "spla_pzgemm_sbs complex, mLocal=100000 n=4096 k=4096, "
                                  "device-resident, beta=0", "dist": "block cyclic 128x128 on 1x1",
                      "split_tolerance": tol},
           "dtype": "c128", "forms": {}}
    for form in forms:
        ctx.set_gemm_mode(_modes(spla)[form])
        run = lambda: spla.pgemm_sbs(mL, n, k, 1.0, dA, mL, dB, k, 0, 0, dist, 0.0, dC, mL, ctx)
        wall, ev = _timed(run, steps, warmup, stream)
        err = normwise_from_samples(_gather(dC, mL, n, rows, cols), ref, mL, n, nAB)
        info = ctx.last_product_info()
        out["forms"][form] = {"tflops": round(flops / wall / 1e9, 3), "ms_per_step": round(wall, 3),
                              "device_ms_per_step": round(ev, 3), "normwise_error": err,
                              "error_samples": len(rows) * len(cols), **_info_keys(info),
                              "phase_ms": _phases(ctx, run)}
        log(f"sbs100k {form}: {out['forms'][form]}")
    ctx.close()
    del dA, dB, dC
    _release()
    return out


def workload_ssb_real_spread(spla, dev, stream, steps, warmup, forms, tol):
    """pdgemm_ssb A^T B real, m = n = 2048, k = 2e5, entries
    uniform[-1,1] * 2^e with e uniform in [-10, 10]; a wider range; and rows with one
    dominant entry, which the split form's bound cannot serve (FP64 fall-back)."""
    import torch

    m = n = 2048
    k = 200000

    def spread(mid, count, lo, hi):
        x = uniform_device(SEED, mid, count, False, dev)
        e = (uniform_device(SEED, mid + 100, count, False, dev) + 1.0) * 0.5 * (hi - lo + 1)
        e = torch.clamp(torch.floor(e) + lo, lo, hi).to(torch.int32)
        return torch.ldexp(x, e)

    def dominant(mid, rows_):
        # one entry of magnitude ~1 per row (column of A), the rest ~1e-40
        x = uniform_device(SEED, mid, k * rows_, False, dev) * 1e-40
        x.view(rows_, k)[:, 0] = 1.0
        return x

    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_stream(stream)
    ctx.set_split_tolerance(tol)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 256,
                                                             256)
    flops = 2.0 * m * n * k
    rows, cols = sampled_rows_cols(m, n, 32, 3)
    out = {"config": {"workload": "spla_pdgemm_ssb A^T B real, m=n=2048 k=200000, entries "
                                  "uniform[-1,1] * 2^e", "dist": "block cyclic 256x256 on 1x1",
                      "split_tolerance": tol},
           "dtype": "f64", "forms": {}}
    cases = (("e in [-10, 10]", lambda mid, r: spread(mid, k * r, -10, 10)),
             ("e in [-300, 300]", lambda mid, r: spread(mid, k * r, -300, 300)),
             ("one dominant entry per row of A (forced fall-back)",
              lambda mid, r: dominant(mid, r) if mid == 11 else spread(mid, k * r, -10, 10)))
    for case, gen in cases:
        dA = gen(11, m)
        dB = gen(12, n)
        dC = torch.zeros(m * n, dtype=torch.float64, device=dev)
        ref = _ssb_reference(dA, dB, k, m, n, rows, cols, conj=False)
        nAB = torch.linalg.vector_norm(dA).item() * torch.linalg.vector_norm(dB).item()
        case_forms = forms if not case.startswith("one") else [f for f in forms if f != "plain"]
        for form in case_forms:
            ctx.set_gemm_mode(_modes(spla)[form])
            run = lambda: spla.pgemm_ssb(m, n, k, spla.Operation.TRANSPOSE, 1.0, dA, k, dB, k, 0.0,
                                         dC, m, 0, 0, dist, ctx)
            wall, ev = _timed(run, steps, warmup, stream)
            err = normwise_from_samples(_gather(dC, m, n, rows, cols), ref, m, n, nAB)
            info = ctx.last_product_info()
            key = f"{form}, {case}"
            out["forms"][key] = {"tflops": round(flops / wall / 1e9, 3),
                                 "ms_per_step": round(wall, 3), "device_ms_per_step": round(ev, 3),
                                 "normwise_error": err, "error_samples": len(rows) * len(cols),
                                 **_info_keys(info), "phase_ms": _phases(ctx, run)}
            log(f"ssb_real_spread {key}: {out['forms'][key]}")
        del dA, dB, dC
        _release()
    sp = out["forms"].get("split, e in [-10, 10]")
    pl = out["forms"].get("plain, e in [-10, 10]")
    if sp and pl:
        out["split_over_plain"] = round(sp["tflops"] / pl["tflops"], 3)
    ctx.close()
    return out


def workload_batch512(spla, dev, stream, steps, warmup, forms, tol):
    """512 uneven complex ring-step blocks (m, n log-uniform in
    [32, 512], k in [4096, 65536], seed 5; odd blocks LOWER fill) through the batch
    API, alpha = 1, beta = 0."""
    import numpy as np
    import torch

    rng = np.random.default_rng(5)
    cnt = 512
    ms_ = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), cnt))).astype(int)
    ns_ = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), cnt))).astype(int)
    ks_ = np.rint(np.exp(rng.uniform(np.log(4096), np.log(65536), cnt))).astype(int)
    offA = np.concatenate([[0], np.cumsum(ks_ * ms_)])
    offB = np.concatenate([[0], np.cumsum(ks_ * ns_)])
    offC = np.concatenate([[0], np.cumsum(ms_ * ns_)])
    dA = uniform_device(SEED, 1, int(offA[-1]), True, dev)
    dB = uniform_device(SEED, 2, int(offB[-1]), True, dev)
    dC = torch.zeros(int(offC[-1]), dtype=torch.complex128, device=dev)
    specs = []
    full = tri = 0.0
    for i in range(cnt):
        mi, ni, ki = int(ms_[i]), int(ns_[i]), int(ks_[i])
        fill = spla.FillMode.LOWER if i % 2 else spla.FillMode.FULL
        specs.append(dict(m=mi, n=ni, k=ki, A=dA.data_ptr() + 16 * int(offA[i]), lda=ki,
                          B=dB.data_ptr() + 16 * int(offB[i]), ldb=ki,
                          C=dC.data_ptr() + 16 * int(offC[i]), ldc=mi, fill=fill))
        full += 8.0 * mi * ni * ki
        req = mi * ni if i % 2 == 0 else sum(min(ni, r + 1) for r in range(mi))
        tri += 8.0 * req * ki
    # accuracy: 8 x 8 sampled entries of every 16th block (required triangle only)
    checks = []
    nAB2 = 0.0
    for i in range(0, cnt, 16):
        mi, ni, ki = int(ms_[i]), int(ns_[i]), int(ks_[i])
        rr, cc = sampled_rows_cols(mi, ni, 8, 100 + i)
        Ai = dA[int(offA[i]):int(offA[i + 1])]
        Bi = dB[int(offB[i]):int(offB[i + 1])]
        checks.append((i, rr, cc, _ssb_reference(Ai, Bi, ki, mi, ni, rr, cc),
                       torch.linalg.vector_norm(Ai).item() * torch.linalg.vector_norm(Bi).item()))
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_stream(stream)
    ctx.set_split_tolerance(tol)
    out = {"config": {"workload": "512 complex ssb ring-step blocks A^H B through the batch API "
                                  "(m, n log-uniform in [32, 512], k in [4096, 65536], seed 5; "
                                  "odd blocks LOWER fill), alpha=1 beta=0",
                      "flops_full_blocks": full, "flops_required_triangle": tri,
                      "split_tolerance": tol},
           "dtype": "c128", "forms": {}}
    torch.cuda.synchronize()
    for form in forms:
        ctx.set_gemm_mode(_modes(spla)[form])
        # the FP64 launch's tile shape, split scale and dispatch from timed candidates
        # (the autotuner); its choice then serves the automatic mode's FP64 launch too
        batch = spla.Batch(ctx, specs, 1.0, 0.0, measure=form == "plain")
        wall, ev = _timed(batch.execute, steps, warmup, stream)
        worst = 0.0
        for i, rr, cc, ref, nAB in checks:
            mi, ni = int(ms_[i]), int(ns_[i])
            got = _gather(dC[int(offC[i]):int(offC[i + 1])], mi, ni, rr, cc)
            keep = torch.ones_like(ref.real, dtype=torch.bool)
            if i % 2:  # LOWER: rows >= cols required
                keep = torch.from_numpy(rr[:, None] >= cc[None, :]).to(dev)
            d = (got - ref).abs()[keep] / nAB
            err = float(torch.sqrt((d * d).mean() * (mi * ni)).item())
            worst = max(worst, err)
        out["forms"][form] = {"tflops_full_blocks": round(full / wall / 1e9, 3),
                              "tflops_required_triangle": round(tri / wall / 1e9, 3),
                              "ms_per_step": round(wall, 3), "device_ms_per_step": round(ev, 3),
                              "worst_block_normwise_error": worst, "blocks_checked": len(checks),
                              "plan": batch.info()}
        batch.close()
        log(f"batch512 {form}: {out['forms'][form]}")
    ctx.close()
    del dA, dB, dC
    _release()
    return out


def workload_ssb2048_c64(spla, dev, stream, steps, warmup, forms, tol):
    """The headline shape in single precision: FP32 matrix cores, FP32 accumulation."""
    import torch

    m, n, k = M, N, K
    dA = uniform_device(SEED, 1, k * m, True, dev).to(torch.complex64)
    dB = uniform_device(SEED, 2, k * n, True, dev).to(torch.complex64)
    dC = torch.zeros(m * n, dtype=torch.complex64, device=dev)
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_stream(stream)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 256,
                                                             256)
    flops = 8.0 * m * n * k
    rows, cols = sampled_rows_cols(m, n, 32, 3)
    ref = _ssb_reference(dA.to(torch.complex128), dB.to(torch.complex128), k, m, n, rows, cols)
    nAB = (torch.linalg.vector_norm(dA.to(torch.complex128)).item() *
           torch.linalg.vector_norm(dB.to(torch.complex128)).item())
    run = lambda: spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, 1.0, dA, k, dB, k, 0.0,
                                 dC, m, 0, 0, dist, ctx)
    wall, ev = _timed(run, steps, warmup, stream)
    got = _gather(dC, m, n, rows, cols).to(torch.complex128)
    err = normwise_from_samples(got, ref, m, n, nAB)
    out = {"config": {"workload": "spla_pcgemm_ssb A^H B complex float m=n=2048 k=200000, "
                                  "device-resident, FP32 accumulation"},
           "dtype": "c64",
           "forms": {"plain": {"tflops": round(flops / wall / 1e9, 3),
                               "ms_per_step": round(wall, 3), "device_ms_per_step": round(ev, 3),
                               "normwise_error": err,
                               "fraction_of_fp32_matrix_peak": round(flops / wall / 1e9 / 157.3, 4),
                               "phase_ms": _phases(ctx, run)}}}
    log(f"ssb2048_c64: {out['forms']['plain']}")
    ctx.close()
    del dA, dB, dC
    _release()
    return out


def workload_ssb_small_mn(spla, dev, stream, steps, warmup, forms, tol):
    """HBM-bound shapes: complex intensity m n / (2 (m + n)) FLOP/B is below the FP64
    ridge 78.6 / 8 = 9.8 FLOP/B for m = n <= 39; plain FP64 kernels (split-K)."""
    import torch

    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_stream(stream)
    ctx.set_gemm_mode(spla.GemmMode.FP64)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 256,
                                                             256)
    out = {"config": {"workload": "spla_pzgemm_ssb A^H B complex, m = n = 16 (k = 8e6) and "
                                  "32 (k = 4e6), device-resident, plain FP64, HBM-bound"},
           "dtype": "f64", "forms": {}}
    for (m, n, k) in ((16, 16, 8_000_000), (32, 32, 4_000_000)):
        dA = uniform_device(SEED, 1, k * m, True, dev)
        dB = uniform_device(SEED, 2, k * n, True, dev)
        dC = uniform_device(SEED, 3, m * n, True, dev)
        C0 = dC.clone()
        flops = 8.0 * m * n * k
        hbm = 16.0 * k * (m + n) + 2 * 16.0 * m * n  # A, B once; C read (beta != 0) and written
        run = lambda: spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, ALPHA, dA, k, dB,
                                     k, BETA, dC, m, 0, 0, dist, ctx)
        dC.copy_(C0)
        run()
        rows, cols = sampled_rows_cols(m, n, 16, 5)
        ref = ALPHA * _ssb_reference(dA, dB, k, m, n, rows, cols) + BETA * _gather(C0, m, n, rows,
                                                                                   cols)
        nAB = torch.linalg.vector_norm(dA).item() * torch.linalg.vector_norm(dB).item()
        err = normwise_from_samples(_gather(dC, m, n, rows, cols), ref, m, n, nAB)
        wall, ev = _timed(run, steps, warmup, stream)
        gbs = hbm / (wall / 1e3) / 1e9
        key = f"m=n={m} k={k}"
        out["forms"][key] = {"tflops": round(flops / wall / 1e9, 3), "ms_per_step": round(wall, 3),
                             "device_ms_per_step": round(ev, 3), "hbm_gb_per_s": round(gbs, 1),
                             "fraction_of_hbm_peak": round(gbs / 8000.0, 4),
                             "intensity_flop_per_byte": round(flops / hbm, 2),
                             "normwise_error": err, "phase_ms": _phases(ctx, run)}
        log(f"ssb_small_mn {key}: {out['forms'][key]}")
        del dA, dB, dC, C0
        _release()
    ctx.close()
    return out


WORKLOAD_FNS = {
    "ssb2048_host": workload_ssb2048_host,
    "sbs100k": workload_sbs100k,
    "ssb_real_spread": workload_ssb_real_spread,
    "batch512": workload_batch512,
    "ssb2048_c64": workload_ssb2048_c64,
    "ssb_small_mn": workload_ssb_small_mn,
}


def copy_calibration(dev) -> None:
    """A 1 GiB device-to-device copy: the reference kernel of the HBM counters
    (1 GiB read + 1 GiB written)."""
    import torch

    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log("copy calibration: 1 GiB read + 1 GiB written")
    del src, dst


def side_workload(args) -> int:
    """One of the other parity workloads as its own JSON line (--workload NAME)."""
    import torch

    import spla_amd as spla

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    steps, warmup = max(1, args.steps), max(0, args.warmup)
    if args.copy_calibration:
        copy_calibration(dev)
    res = WORKLOAD_FNS[args.workload](spla, dev, stream, steps, warmup, args.forms.split(","),
                                      args.split_tolerance)
    out = {"metric": METRIC, "unit": "TFLOP/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "higher_is_better": True, "data": "synthetic (splitmix64 seed 42)", **res}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--forms", default="plain,split", help="comma list of: plain, split, auto")
    ap.add_argument("--split-tolerance", type=float, default=1e-13,
                    help="normwise error bound the split form is planned for (the metric's "
                         "required accuracy; the library default is 1e-14)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the other BASELINE configurations after the headline")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device (reference kernel for HBM counters)")
    ap.add_argument("--workload", default="ssb2048", choices=WORKLOADS,
                    help="ssb2048 (default, the reported line) or one of the other parity "
                         "workloads, each printed as its own JSON line")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU (single rank); --gpus must be 1",
              file=sys.stderr)
        return 2
    if args.workload != "ssb2048":
        return side_workload(args)

    import torch

    import spla_amd as spla

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    m, n, k = M, N, K
    log(f"generating A {k}x{m}, B {k}x{n} complex on {dev}")
    dA = uniform_device(SEED, 1, k * m, True, dev)
    dB = uniform_device(SEED, 2, k * n, True, dev)
    dC0 = uniform_device(SEED, 3, m * n, True, dev)
    torch.cuda.synchronize()

    ctx = spla.Context(spla.ProcessingUnit.GPU)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 256,
                                                             256)
    # library initialisation (code objects, streams) on a tiny product
    tiny = torch.zeros(64 * 64, dtype=torch.complex128, device=dev)
    spla.pgemm_ssb(64, 64, 128, spla.Operation.CONJ_TRANSPOSE, ALPHA, dA, k, dB, k, BETA, tiny,
                   64, 0, 0, dist, ctx)

    if args.copy_calibration:
        copy_calibration(dev)

    flops = 8.0 * m * n * k
    modes = {"plain": spla.GemmMode.FP64, "split": spla.GemmMode.SPLIT_INT8,
             "auto": spla.GemmMode.AUTO}
    forms = {}
    ctx.set_split_tolerance(args.split_tolerance)
    for form in args.forms.split(","):
        ctx.set_gemm_mode(modes[form])
        ctx.set_timing(True)
        dC = dC0.clone()
        for w in range(args.warmup):
            spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, ALPHA, dA, k, dB, k, BETA, dC,
                           m, 0, 0, dist, ctx)
            log(f"{form}: warmup step {w + 1}/{args.warmup} done")
        dC.copy_(dC0)
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
        first_result = None
        phase_sum = {}
        t0 = time.perf_counter()
        for s in range(args.steps):
            evs[2 * s].record(stream)
            spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, ALPHA, dA, k, dB, k, BETA, dC,
                           m, 0, 0, dist, ctx)
            evs[2 * s + 1].record(stream)
            for name, v in ctx.last_timing().items():
                phase_sum[name] = phase_sum.get(name, 0.0) + v
            if s == 0 and args.steps > 1:
                first_result = dC.clone()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        log(f"{form}: {args.steps} steps done")
        kernel_ms = sum(evs[2 * s].elapsed_time(evs[2 * s + 1]) for s in range(args.steps))
        kernel_ms /= args.steps
        info = ctx.last_product_info()
        res = first_result if first_result is not None else dC
        entry = {"tflops": round(flops / wall / 1e12, 3), "ms_per_step": round(wall * 1e3, 3),
                 "device_ms_per_step": round(kernel_ms, 3),
                 "phase_ms_per_step": {p: round(v / args.steps, 4)
                                       for p, v in phase_sum.items() if v > 0}}
        if form in ("split", "auto"):
            entry["split_used"] = info["split"]
            entry["moduli"] = info["moduli"]
            entry["bits"] = [info["bits_a"], info["bits_b"]]
            entry["error_bound"] = info["bound"]
            entry["fallback_reason"] = info["reason"]
        if not args.no_check:
            err, cnt = sampled_error(dA, dB, dC0, res, k, m, n)
            entry["normwise_error"] = err
            entry["error_samples"] = cnt
            entry["accurate"] = err <= 1e-13
            log(f"{form}: {entry['tflops']} TFLOP/s, sampled normwise error {err:.2e}")
        forms[form] = entry
        ctx.set_timing(False)

    ok = {f: e for f, e in forms.items() if e.get("accurate", True)}
    best_form = max(ok, key=lambda f: ok[f]["tflops"]) if ok else "plain"
    best = forms[best_form]
    if best_form in ("split", "auto") and best.get("split_used"):
        # dominant kernel: the batched INT8 GEMM; algorithmic ops = 2 m n k per product
        # (3 complex sub-products per modulus; the zero padding of k is not counted)
        ops = 2.0 * m * n * k * 3 * best["moduli"]
        gemm_ms = best["phase_ms_per_step"].get("split_int8_gemm", float("nan"))
        achieved = ops / (gemm_ms / 1e3) / 1e12
        form = ("v_mfma_i32_32x32x32_i8" if os.environ.get("SPLA_AMD_I8_MFMA") == "32"
                else "v_mfma_i32_16x16x64_i8")
        roofline = {"bound": "mfma", "kernel": f"i8_gemm_kernel ({form})",
                    "achieved": round(achieved, 2), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                    "frac": round(achieved / INT8_PEAK_TOPS, 4),
                    # counters were taken on the plan with PROFILED_MODULI moduli
                    "traffic": (HBM_TRAFFIC_BYTES["split"] if best["moduli"] == PROFILED_MODULI
                                else None)}
        dtype = "i8"
    else:
        gemm_ms = best["phase_ms_per_step"].get("gemm_fp64", best["device_ms_per_step"])
        achieved = flops / (gemm_ms / 1e3) / 1e12
        roofline = {"bound": "mfma", "kernel": "gemm_f64_kernel (v_mfma_f64_16x16x4_f64)",
                    "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                    "traffic": HBM_TRAFFIC_BYTES["plain"]}
        dtype = "f64"
    line = {
        "metric": METRIC,
        "value": best["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (splitmix64 seed 42, uniform [-1,1) per component), device-resident",
        "config": {"workload": "spla_pzgemm_ssb A^H B complex double, 1 rank (MPI_COMM_SELF)",
                   "m": m, "n": n, "k": k, "alpha": 1.0, "beta": 0.5,
                   "distC": "block cyclic 256x256 on 1x1", "form": best_form,
                   "split_tolerance": args.split_tolerance},
        "forms": forms,
        "roofline": roofline,
    }
    ctx.close()
    del dA, dB, dC0, dC
    _release()
    if not args.no_workloads:
        # every other BASELINE configuration, measured in the same run (fewer steps)
        line["workloads"] = {}
        ws, ww = max(1, min(args.steps, 3)), 1
        for name in DEFAULT_SIDE:
            log(f"workload {name}")
            t0 = time.perf_counter()
            try:
                res = WORKLOAD_FNS[name](spla, dev, stream, ws, ww, SIDE_FORMS[name],
                                         args.split_tolerance)
            except Exception as exc:  # recorded, the headline line is still printed
                res = {"error": f"{type(exc).__name__}: {exc}"}
                log(f"workload {name} failed: {res['error']}")
                _release()
            res["steps"], res["warmup"] = ws, ww
            res["wall_s"] = round(time.perf_counter() - t0, 1)
            line["workloads"][name] = res
    if not args.no_cpu_baseline:
        log("CPU baseline (host BLAS, k-truncated sample)")
        line["cpu_baseline"] = cpu_baseline(m, n, k)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
