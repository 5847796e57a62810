"""Autotuner of the kernel launch shapes.

Every size n has up to four compiled kernel shapes (lanes per matrix,
lane-columns, waves per SIMD the register allocation allows;
``candidate_shape`` in csrc/lup_kernel.hpp).  This times each one on the
device on a seeded N(0,1) batch of about ``target_mib`` MiB, checks that it
returns the same determinants bit for bit as shape 0, and writes the fastest
per (policy, n) to a JSON table.  The Python binding applies the table when
it loads the library (lucore.SHAPE_TABLE); C and C++ callers apply it with
lufqhe_load_shape_table, which LUCoreGPU::CALL does on first use.

    python lion-codes_amd/autotune.py [--out PATH] [--ns 8,16,32] [--reps 5] [--count N]
"""
from __future__ import annotations

import argparse
import importlib
import json
import sys
import time
from pathlib import Path

import torch

_PKG = Path(__file__).resolve().parent
if str(_PKG.parent) not in sys.path:
    sys.path.insert(0, str(_PKG.parent))


def time_shape(lc, A, res, policy: str, n: int, shape: int, reps: int) -> float:
    """Mean kernel time (ms) of `reps` launches with HIP events on the stream."""
    assert lc.set_shape(policy, "fused", n, shape)
    stream = torch.cuda.current_stream()
    lc.det_batched(A, policy=policy, out=res)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        lc.det_batched(A, policy=policy, out=res)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / reps


def tune(policies=("crout_ref", "abs"), ns=range(1, 65), target_mib: int = 256, reps: int = 5,
         seed: int = 4242, log=print, count: int = 0) -> dict:
    lc = importlib.import_module("lion-codes_amd")
    wl = importlib.import_module("lion-codes_amd.workloads")
    lc.load()
    dev = torch.device("cuda:0")
    # keep the device busy for a while first: its clocks ramp up over the
    # first tens of milliseconds of work, which would favour later shapes
    A = wl.gaussian(65536, 32, seed, device=dev)
    res = lc.alloc_result(65536, 32, dev)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.3:
        for _ in range(10):
            lc.det_batched(A, out=res)
        torch.cuda.synchronize()
    del A, res
    entries = []
    for policy in policies:
        for n in ns:
            cnt = count or int(min(262144, max(4096, (target_mib << 20) // (16 * n * n))))
            A = wl.gaussian(cnt, n, seed + n, device=dev)
            res = lc.alloc_result(count, n, dev)
            ref = None
            times = []
            for shape in range(lc.shape_count(n)):
                t = time_shape(lc, A, res, policy, n, shape, reps)
                mant = res.mant.clone()
                if ref is None:
                    ref = mant
                elif not torch.equal(torch.view_as_real(ref).view(torch.int64),
                                     torch.view_as_real(mant).view(torch.int64)):
                    raise RuntimeError(f"shape {shape} of n={n} ({policy}) differs from shape 0")
                times.append(t)
            best = min(range(len(times)), key=times.__getitem__)
            lc.set_shape(policy, "fused", n, best)
            lanes, cols, waves = lc.shape_info(n, best)
            entries.append({"policy": policy, "n": n, "shape": best, "lanes": lanes,
                            "lane_cols": cols, "waves": waves, "count": cnt,
                            "us_per_launch": [round(t * 1e3, 2) for t in times],
                            "ns_per_det": round(times[best] * 1e6 / cnt, 3)})
            log(f"{policy:9s} n={n:2d} shape={best} (L={lanes}, CG={cols}, waves={waves}) "
                f"times(us)={[round(t * 1e3, 1) for t in times]}")
            del A, res
    props = torch.cuda.get_device_properties(0)
    return {"device": props.name, "arch": getattr(props, "gcnArchName", ""),
            "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S"), "batch_mib": target_mib,
            "reps": reps, "entries": entries}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(_PKG / "shapes_gfx950.json"))
    ap.add_argument("--ns", default="1-64")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--count", type=int, default=0,
                    help="matrices per size (default: about --mib MiB, at least 4096)")
    args = ap.parse_args()
    if "-" in args.ns:
        lo, hi = map(int, args.ns.split("-"))
        ns = range(lo, hi + 1)
    else:
        ns = [int(x) for x in args.ns.split(",")]
    table = tune(ns=ns, target_mib=args.mib, reps=args.reps, log=lambda m: print(m, flush=True),
                 count=args.count)
    Path(args.out).write_text(json.dumps(table, indent=1))
    print(f"wrote {args.out} ({len(table['entries'])} entries)", flush=True)


if __name__ == "__main__":
    main()
