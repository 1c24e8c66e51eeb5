"""Benchmark of the local GEMM engine on one MI355X.

Workload (one step): spla_pzgemm_ssb on one rank (MPI_COMM_SELF), complex
double, C <- alpha * A^H * B + beta * C with A (k = 200000 x m = 2048) and
B (k x n = 2048), C 2048 x 2048 block cyclic with 256 x 256 blocks on a 1 x 1
grid, alpha = 1, beta = 0.5; A, B and C hold values uniform in [-1, 1) per
component (splitmix64, seed 42, matrix ids 1/2/3) and are resident in HBM
when the timed region starts.

Forms: "plain" (FP64 MFMA kernels) and "split" (FP64-accurate split precision:
Ozaki scheme II on the INT8 matrix cores, falling back to FP64 when its error
bound cannot be met).

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
# copy (reads x 2.0, writes x 1.0). They change only when the kernels change.
HBM_TRAFFIC_BYTES = {"split": 67.06e9 + 1.51e9, "plain": 244.86e9 + 0.07e9}
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

def sampled_error(dA, dB, dC0, dC, k, m, n, samples=48, seed=3):
    """Normwise error estimate from sampled entries against the double-double oracle."""
    import numpy as np
    import torch

    import oracle

    rng = np.random.default_rng(seed)
    rows = rng.integers(0, m, samples)
    cols = rng.integers(0, n, samples)
    ur, rinv = np.unique(rows, return_inverse=True)
    uc, cinv = np.unique(cols, return_inverse=True)
    A_cols = dA.view(m, k)[torch.from_numpy(ur).to(dA.device)].cpu().numpy()  # (len(ur), k)
    B_cols = dB.view(n, k)[torch.from_numpy(uc).to(dB.device)].cpu().numpy()
    C0 = dC0.view(n, m).cpu().numpy()  # column-major: C0[j, i]
    got = dC.view(n, m).cpu().numpy()
    ref = np.empty(samples, dtype=np.complex128)
    for s in range(samples):
        a = np.ascontiguousarray(A_cols[rinv[s]])
        b = np.ascontiguousarray(B_cols[cinv[s]])
        c0 = np.array([C0[cols[s], rows[s]]], dtype=np.complex128)
        ref[s] = oracle.gemm_entries(2, 0, k, ALPHA, a, k, b, k, BETA, c0, 1, [0], [0])[0]
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
                   f"host cores visible {cores}"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--forms", default="plain,split", help="comma list of: plain, split")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device (reference kernel for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU (single rank); --gpus must be 1",
              file=sys.stderr)
        return 2

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
    modes = {"plain": spla.GemmMode.FP64, "split": spla.GemmMode.SPLIT_INT8}
    forms = {}
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
        if form == "split":
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
    if best_form == "split" and best.get("split_used"):
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
                   "distC": "block cyclic 256x256 on 1x1", "form": best_form},
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
