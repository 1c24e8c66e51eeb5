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
whose measured normwise error ||C - C_dd|| / (||A|| ||B||) (sampled entries
against the double-double oracle) is <= 1e-13. Every form run is reported.

roofline: the dominant kernel of that form (FP64 GEMM kernel, or the batched
INT8 GEMM of the split form), its duration from HIP events the library records
on its launch stream (per-phase timing of the context), against MI355X's
published peak (FP64 matrix 78.6 TFLOP/s; INT8 dense 5.0 POP/s).

cpu_baseline: SPLA's CPU path (one host-BLAS zgemm, as in
src/gemm/gemm_host.cpp:55-72) on a k-truncated sample of the same workload,
timed on this host's cores.
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
HBM_TRAFFIC_BYTES = {"split": 51.66e9 + 1.41e9, "plain": 116.67e9 + 0.07e9}
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
    """Error-free product x * y = p + e (Veltkamp splitting, no FMA needed)."""
    import numpy as np

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
    errors of every addition are kept and added at the end."""
    import numpy as np

    v = np.asarray(v, dtype=np.float64)
    errs = np.zeros(v.shape[:-1])
    while v.shape[-1] > 1:
        if v.shape[-1] % 2:
            v = np.concatenate([v, np.zeros(v.shape[:-1] + (1,))], axis=-1)
        a, b = v[..., 0::2], v[..., 1::2]
        s = a + b
        bb = s - a
        errs = errs + ((a - (s - bb)) + (b - bb)).sum(axis=-1)
        v = s
    return v[..., 0] + errs


def _accurate_dot(x, y):
    """sum_l x[..., l] * y[..., l] to about double-double accuracy."""
    import numpy as np

    p, e = _two_prod(x, y)
    return _accurate_sum(np.concatenate([p, e], axis=-1))


def sampled_error(dA, dB, dC0, dC, k, m, n, samples=48, seed=3):
    """Normwise error estimate ||C - C_ref|| / (||A|| ||B||) from sampled entries;
    C_ref = alpha A^H B + beta C0 with error-free products and compensated sums."""
    import numpy as np
    import torch

    rng = np.random.default_rng(seed)
    rows = rng.integers(0, m, samples)
    cols = rng.integers(0, n, samples)
    ur, rinv = np.unique(rows, return_inverse=True)
    uc, cinv = np.unique(cols, return_inverse=True)
    A_cols = dA.view(m, k)[torch.from_numpy(ur).to(dA.device)].cpu().numpy()  # (len(ur), k)
    B_cols = dB.view(n, k)[torch.from_numpy(uc).to(dB.device)].cpu().numpy()
    a = A_cols[rinv]
    b = B_cols[cinv]
    ar, ai, br, bi = a.real, a.imag, b.real, b.imag
    # conj(a) . b = (ar br + ai bi) + i (ar bi - ai br)
    re = _accurate_dot(np.concatenate([ar, ai], axis=-1), np.concatenate([br, bi], axis=-1))
    im = _accurate_dot(np.concatenate([ar, -ai], axis=-1), np.concatenate([bi, br], axis=-1))
    C0 = dC0.view(n, m).cpu().numpy()  # column-major: C0[j, i]
    got = dC.view(n, m).cpu().numpy()
    ref = ALPHA * (re + 1j * im) + BETA * C0[cols, rows]
    diff = got[cols, rows] - ref
    est = np.sqrt(np.mean(np.abs(diff) ** 2) * m * n)
    nA = torch.linalg.vector_norm(dA).item()
    nB = torch.linalg.vector_norm(dB).item()
    return float(est / (nA * nB)), samples


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
    return {
        "value": round(rate, 4),
        "unit": "TFLOP/s",
        "cores": threads if threads and threads > 0 else cores,
        "kind": "port",
        "sample": (f"{info['vendor']} cblas_zgemm('C','N') m=n={m}, k={k2} (of {k_full}); "
                   f"{t2:.2f} s; extrapolated full-k step {t2 * k_full / k2:.1f} s; "
                   f"all threads at k={k1}: {8.0 * m * n * k1 / t1 / 1e12:.4f} TFLOP/s{one}; "
                   f"host cores visible {cores}"),
    }


WORKLOADS = ("ssb2048", "sbs100k", "ssb_real_spread", "batch512", "ssb2048_host",
             "ssb2048_c64", "ssb_small_mn")


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


def side_workload(args) -> int:
    """The other parity workloads: sbs with a tall local A, a real ssb product with a
    spread of exponents (and a range that forces the FP64 fall-back), 512 uneven
    ring-step blocks through the batch API, and the headline product with
    host-resident operands (PCIe-inclusive). One JSON line each."""
    import numpy as np
    import torch

    import spla_amd as spla

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    ctx = spla.Context(spla.ProcessingUnit.GPU)
    ctx.set_stream(stream)
    modes = {"plain": spla.GemmMode.FP64, "split": spla.GemmMode.SPLIT_INT8,
             "auto": spla.GemmMode.AUTO}
    steps, warmup = max(1, min(args.steps, 5)), max(1, min(args.warmup, 1))
    out = {"metric": METRIC, "unit": "TFLOP/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "higher_is_better": True, "data": "synthetic (splitmix64 seed 42)", "forms": {}}
    wl = args.workload

    if wl == "sbs100k":
        mL, n, k = 100000, 4096, 4096
        dA = uniform_device(SEED, 1, mL * k, True, dev)
        dB = uniform_device(SEED, 2, k * n, True, dev)
        dC = torch.zeros(mL * n, dtype=torch.complex128, device=dev)
        dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1,
                                                                 128, 128)
        flops = 8.0 * mL * n * k
        rows = np.random.default_rng(3).integers(0, mL, 32)
        cols = np.random.default_rng(4).integers(0, n, 32)
        a = dA.view(k, mL)[:, torch.from_numpy(rows).to(dev)].T.cpu().numpy()  # rows of A
        b = dB.view(n, k)[torch.from_numpy(cols).to(dev)].cpu().numpy()
        re = _accurate_dot(np.concatenate([a.real, -a.imag], -1), np.concatenate([b.real, b.imag], -1))
        im = _accurate_dot(np.concatenate([a.real, a.imag], -1), np.concatenate([b.imag, b.real], -1))
        ref = re + 1j * im
        nAB = torch.linalg.vector_norm(dA).item() * torch.linalg.vector_norm(dB).item()
        for form in args.forms.split(","):
            ctx.set_gemm_mode(modes[form])
            run = lambda: spla.pgemm_sbs(mL, n, k, 1.0, dA, mL, dB, k, 0, 0, dist, 0.0, dC, mL, ctx)
            wall, ev = _timed(run, steps, warmup, stream)
            got = dC.view(n, mL).cpu().numpy()[cols, rows]
            err = float(np.sqrt(np.mean(np.abs((got - ref) / nAB) ** 2) * mL * n))
            info = ctx.last_product_info()
            out["forms"][form] = {"tflops": round(flops / wall / 1e9, 3), "ms_per_step": round(wall, 3),
                                  "device_ms_per_step": round(ev, 3), "normwise_error": err,
                                  "split_used": info["split"], "moduli": info["moduli"],
                                  "phase_ms": _phases(ctx, run)}
            log(f"sbs100k {form}: {out['forms'][form]}")
        out["config"] = {"workload": "spla_pzgemm_sbs complex, mLocal=100000 n=4096 k=4096, "
                                     "device-resident, beta=0", "dist": "block cyclic 128x128 on 1x1"}
        out["dtype"] = "c128"

    elif wl == "ssb_real_spread":
        m = n = 2048
        k = 200000

        def spread(mid, count, lo, hi):
            x = uniform_device(SEED, mid, count, False, dev)
            e = (uniform_device(SEED, mid + 100, count, False, dev) + 1.0) * 0.5 * (hi - lo + 1)
            e = torch.clamp(torch.floor(e) + lo, lo, hi).to(torch.int32)
            return torch.ldexp(x, e)

        flops = 2.0 * m * n * k
        rows = np.random.default_rng(3).integers(0, m, 32)
        cols = np.random.default_rng(4).integers(0, n, 32)
        for case, (lo, hi) in (("e in [-10, 10]", (-10, 10)), ("e in [-300, 300]", (-300, 300))):
            dA = spread(11, k * m, lo, hi)
            dB = spread(12, k * n, lo, hi)
            dC = torch.zeros(m * n, dtype=torch.float64, device=dev)
            dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1,
                                                                     256, 256)
            a = dA.view(m, k)[torch.from_numpy(rows).to(dev)].cpu().numpy()
            b = dB.view(n, k)[torch.from_numpy(cols).to(dev)].cpu().numpy()
            ref = _accurate_dot(a, b)
            nAB = torch.linalg.vector_norm(dA).item() * torch.linalg.vector_norm(dB).item()
            for form in args.forms.split(","):
                ctx.set_gemm_mode(modes[form])
                run = lambda: spla.pgemm_ssb(m, n, k, spla.Operation.TRANSPOSE, 1.0, dA, k, dB, k,
                                             0.0, dC, m, 0, 0, dist, ctx)
                wall, ev = _timed(run, steps, warmup, stream)
                got = dC.view(n, m).cpu().numpy()[cols, rows]
                err = float(np.sqrt(np.mean(((got - ref) / nAB) ** 2) * m * n))
                info = ctx.last_product_info()
                out["forms"][f"{form}, {case}"] = {
                    "tflops": round(flops / wall / 1e9, 3), "ms_per_step": round(wall, 3),
                    "device_ms_per_step": round(ev, 3), "normwise_error": err,
                    "split_used": info["split"], "fallback_reason": info["reason"],
                    "moduli": info["moduli"], "bits": [info["bits_a"], info["bits_b"]],
                    "phase_ms": _phases(ctx, run)}
                log(f"ssb_real_spread {form} {case}: {out['forms'][f'{form}, {case}']}")
            del dA, dB, dC
        out["config"] = {"workload": "spla_pdgemm_ssb A^T B real, m=n=2048 k=200000, entries "
                                     "uniform[-1,1] * 2^e", "dist": "block cyclic 256x256 on 1x1"}
        out["dtype"] = "f64"

    elif wl == "batch512":
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
        torch.cuda.synchronize()
        for form in args.forms.split(","):
            ctx.set_gemm_mode(modes[form])
            batch = spla.Batch(ctx, specs, 1.0, 0.0)
            wall, ev = _timed(batch.execute, steps, warmup, stream)
            out["forms"][form] = {"tflops_full_blocks": round(full / wall / 1e9, 3),
                                  "tflops_required_triangle": round(tri / wall / 1e9, 3),
                                  "ms_per_step": round(wall, 3), "device_ms_per_step": round(ev, 3),
                                  "plan": batch.info()}
            batch.close()
            log(f"batch512 {form}: {out['forms'][form]}")
        out["config"] = {"workload": "512 complex ssb ring-step blocks A^H B through the batch API "
                                     "(m, n log-uniform in [32, 512], k in [4096, 65536], seed 5; "
                                     "odd blocks LOWER fill), alpha=1 beta=0",
                         "flops_full_blocks": full, "flops_required_triangle": tri}
        out["dtype"] = "c128"

    elif wl == "ssb2048_host":
        m, n, k = M, N, K
        hA = uniform_device(SEED, 1, k * m, True, dev).cpu().pin_memory()
        hB = uniform_device(SEED, 2, k * n, True, dev).cpu().pin_memory()
        hC = uniform_device(SEED, 3, m * n, True, dev).cpu().pin_memory()
        dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1,
                                                                 256, 256)
        flops = 8.0 * m * n * k
        for form in args.forms.split(","):
            ctx.set_gemm_mode(modes[form])
            run = lambda: spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, ALPHA, hA, k, hB, k,
                                         BETA, hC, m, 0, 0, dist, ctx)
            wall, ev = _timed(run, steps, warmup, stream)
            info = ctx.last_product_info()
            out["forms"][form] = {"tflops_pcie_inclusive": round(flops / wall / 1e9, 3),
                                  "ms_per_step": round(wall, 3), "split_used": info["split"]}
            log(f"ssb2048_host {form}: {out['forms'][form]}")
        out["config"] = {"workload": "spla_pzgemm_ssb A^H B complex m=n=2048 k=200000, A, B, C in "
                                     "pinned host memory (PCIe-inclusive; not the reported value)"}
        out["dtype"] = "c128"

    elif wl == "ssb2048_c64":
        # the headline shape in single precision: FP32 matrix cores, FP32 accumulation
        m, n, k = M, N, K
        dA = uniform_device(SEED, 1, k * m, True, dev).to(torch.complex64)
        dB = uniform_device(SEED, 2, k * n, True, dev).to(torch.complex64)
        dC = torch.zeros(m * n, dtype=torch.complex64, device=dev)
        dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1,
                                                                 256, 256)
        flops = 8.0 * m * n * k
        rows = np.random.default_rng(3).integers(0, m, 32)
        cols = np.random.default_rng(4).integers(0, n, 32)
        a = dA.view(m, k)[torch.from_numpy(rows).to(dev)].cpu().numpy().astype(np.complex128)
        b = dB.view(n, k)[torch.from_numpy(cols).to(dev)].cpu().numpy().astype(np.complex128)
        re = _accurate_dot(np.concatenate([a.real, a.imag], -1), np.concatenate([b.real, b.imag], -1))
        im = _accurate_dot(np.concatenate([a.real, -a.imag], -1), np.concatenate([b.imag, b.real], -1))
        ref = re + 1j * im  # conj(a) . b
        nAB = (torch.linalg.vector_norm(dA.to(torch.complex128)).item() *
               torch.linalg.vector_norm(dB.to(torch.complex128)).item())
        ctx.set_gemm_mode(spla.GemmMode.FP64)
        run = lambda: spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, 1.0, dA, k, dB, k,
                                     0.0, dC, m, 0, 0, dist, ctx)
        wall, ev = _timed(run, steps, warmup, stream)
        got = dC.view(n, m).cpu().numpy().astype(np.complex128)[cols, rows]
        err = float(np.sqrt(np.mean(np.abs((got - ref) / nAB) ** 2) * m * n))
        out["forms"]["plain"] = {"tflops": round(flops / wall / 1e9, 3), "ms_per_step": round(wall, 3),
                                 "device_ms_per_step": round(ev, 3), "normwise_error": err,
                                 "fraction_of_fp32_matrix_peak": round(flops / wall / 1e9 / 157.3, 4),
                                 "phase_ms": _phases(ctx, run)}
        log(f"ssb2048_c64: {out['forms']['plain']}")
        out["config"] = {"workload": "spla_pcgemm_ssb A^H B complex float m=n=2048 k=200000, "
                                     "device-resident, FP32 accumulation"}
        out["dtype"] = "c64"

    elif wl == "ssb_small_mn":
        # HBM-bound shapes: complex intensity m n / (2 (m + n)) FLOP/B is below the FP64
        # ridge 78.6 / 8 = 9.8 FLOP/B for m = n <= 39; plain FP64 kernels (split-K)
        dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1,
                                                                 256, 256)
        ctx.set_gemm_mode(spla.GemmMode.FP64)
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
            rows = np.random.default_rng(5).integers(0, m, 8)
            cols = np.random.default_rng(6).integers(0, n, 8)
            a = dA.view(m, k)[torch.from_numpy(rows).to(dev)].cpu().numpy()
            b = dB.view(n, k)[torch.from_numpy(cols).to(dev)].cpu().numpy()
            re = _accurate_dot(np.concatenate([a.real, a.imag], -1),
                               np.concatenate([b.real, b.imag], -1))
            im = _accurate_dot(np.concatenate([a.real, -a.imag], -1),
                               np.concatenate([b.imag, b.real], -1))
            c0 = C0.view(n, m).cpu().numpy()[cols, rows]
            ref = ALPHA * (re + 1j * im) + BETA * c0
            got = dC.view(n, m).cpu().numpy()[cols, rows]
            nAB = (torch.linalg.vector_norm(dA).item() * torch.linalg.vector_norm(dB).item())
            err = float(np.sqrt(np.mean(np.abs((got - ref) / nAB) ** 2) * m * n))
            wall, ev = _timed(run, steps, warmup, stream)
            gbs = hbm / (wall / 1e3) / 1e9
            key = f"m=n={m} k={k}"
            out["forms"][key] = {"tflops": round(flops / wall / 1e9, 3), "ms_per_step": round(wall, 3),
                                 "device_ms_per_step": round(ev, 3),
                                 "hbm_gb_per_s": round(gbs, 1),
                                 "fraction_of_hbm_peak": round(gbs / 8000.0, 4),
                                 "intensity_flop_per_byte": round(flops / hbm, 2),
                                 "normwise_error": err, "phase_ms": _phases(ctx, run)}
            log(f"ssb_small_mn {key}: {out['forms'][key]}")
            del dA, dB, dC, C0
        out["config"] = {"workload": "spla_pzgemm_ssb A^H B complex, m = n = 16 (k = 8e6) and "
                                     "32 (k = 4e6), device-resident, plain FP64, HBM-bound"}
        out["dtype"] = "f64"

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
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del src, dst

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
        roofline = {"bound": "mfma", "kernel": "i8_gemm_kernel (v_mfma_i32_32x32x32_i8)",
                    "achieved": round(achieved, 2), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                    "frac": round(achieved / INT8_PEAK_TOPS, 4),
                    "traffic": HBM_TRAFFIC_BYTES["split"]}
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
    if not args.no_cpu_baseline:
        log("CPU baseline (host BLAS, k-truncated sample)")
        line["cpu_baseline"] = cpu_baseline(m, n, k)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
