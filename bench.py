"""Benchmark of the batched LUP-determinant hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one pass of the hot path over one batch: one plain launch of the
n=32 kernel over 65536 complex128 matrices (re,im ~ N(0,1), seeded), inputs
resident in HBM, outputs mant/exp/log|det|/phase/info per matrix.  Prints one
JSON line (metric, value = det/s, roofline of the kernel measured with HIP
events on its stream, CPU baseline of the restated reference Crout timed on
this host's cores, accuracy against the long-double truth).
"""
from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("determinants/sec at n=32 complex128 and mixed n=8..64; "
          "p50/p99 latency at 1e-12 rel err")
N, BATCH, SEED = 32, 65536, 20250101
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
OUT_BYTES = 16 + 4 + 8 + 8 + 4   # mant, exp, logabs, phase, info
# HBM bytes one launch of the n=32 kernel moves, measured with rocprofv3
# (FETCH_SIZE and WRITE_SIZE, one counter per run, each scaled by what the
# same counter reads for a plain 1 GiB device copy: x2.0 and x1.0).
MEASURED_TRAFFIC_N32 = 1.0772e9


def det_bytes(n: int) -> int:
    return 16 * n * n + OUT_BYTES


def det_flops(n: int) -> float:
    # LAPACK zgetrf count (6 muls + 2 adds) plus the complex det product
    return 8.0 * n ** 3 / 3.0 - n * n + 13.0 * n / 3.0 + 6.0 * (n - 1)


def host_cores() -> int:
    try:
        c = len(os.sched_getaffinity(0))
    except AttributeError:
        c = os.cpu_count() or 1
    return max(1, min(16, c))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A_host: np.ndarray, cores: int) -> dict:
    """Restated reference CPU Crout (FP64, OpenMP over matrices) on a bounded
    sample of the same batch; median of 5 timed repetitions after a warm-up."""
    from oracle import oracle
    oracle.crout_ref(A_host[:256], nthreads=cores)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        oracle.crout_ref(A_host, nthreads=cores)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    one = A_host[:2048]
    t1 = time.perf_counter()
    oracle.crout_ref(one, nthreads=1)
    t1 = time.perf_counter() - t1
    return {"value": len(A_host) / t, "unit": "det/s", "cores": cores, "kind": "port",
            "sample": f"first {len(A_host)} matrices of the n=32 batch, median of 5 runs; "
                      f"1 thread: {len(one) / t1:.0f} det/s on 2048",
            "cpu_model": cpu_model()}


def accuracy(lc, A, res, count: int = 512) -> dict:
    """|det/det_truth - 1| against the long-double oracle on a sample."""
    import torch
    from oracle import oracle
    idx = torch.arange(0, BATCH, BATCH // count, device=A.device)[:count]
    As = A[idx].cpu().numpy()
    t = oracle.truth(As)
    la = res.logabs[idx].cpu().numpy()
    ph = res.phase[idx].cpu().numpy()
    e = np.abs(np.expm1((la - t["logabs"]) + 1j * np.angle(np.exp(1j * (ph - t["phase"])))))
    return {"vs": "long-double LUP truth", "sample": int(count), "median": float(np.median(e)),
            "p99": float(np.quantile(e, 0.99)), "max": float(e.max()),
            "frac_le_1e-12": float(np.mean(e <= 1e-12))}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; a batch shards with no "
                          "communication, see DESIGN.md -- run with --gpus 1"}))
        sys.exit(2)

    import torch
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 8, "launches": 3}))
        return
    lc = importlib.import_module("lion-codes_amd")
    wl = importlib.import_module("lion-codes_amd.workloads")
    dev = torch.device("cuda:0")
    lc.load()

    A = wl.gaussian(BATCH, N, SEED, device=dev)
    res = lc.alloc_result(BATCH, N, dev)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        lc.det_batched(A, out=res)
    torch.cuda.synchronize()

    # timed region: K steps, bracketed by device synchronisation
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        lc.det_batched(A, out=res)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in ev]))
    assert int((res.info != 0).sum()) == 0

    value = BATCH * args.steps / wall
    bytes_launch = det_bytes(N) * BATCH
    flops_launch = det_flops(N) * BATCH
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "det/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic",
        "config": {"workload": f"uniform batch: {BATCH} complex128 {N}x{N}, re,im ~ N(0,1), "
                               f"one plain launch (config 2)", "batch": BATCH, "n": N,
                   "pivot_policy": "crout_ref", "arith": "fused", "seed": SEED},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC_N32,
                     "kernel_ms": kern_ms,
                     "fp64_tflops": flops_launch / (kern_ms * 1e-3) / 1e12,
                     "fp64_frac": flops_launch / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                     "bytes_per_det": det_bytes(N), "flops_per_det": det_flops(N)},
    }
    out["accuracy"] = accuracy(lc, A, res)
    if not args.no_cpu:
        A_host = A[:16384].cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(A_host, host_cores())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
