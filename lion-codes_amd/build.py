"""Build recipe for the engine's native libraries.

    python -m lion-codes_amd.build            (or build() from Python)

* ``lion-codes_amd/liblufqhe.so`` -- the HIP library (gfx950): kernels, the C
  ABI of ``include/lufqhe.h`` and the C++ API of ``include/lucore_gpu.hpp``.
* ``oracle/liblufqhe_oracle.so`` -- the CPU oracle (gcc, OpenMP), built here
  because the tests and the benchmark's CPU baseline need it; the product
  library never links it.

Kernels are instantiated per (variant, n); ``lup_inst.hip`` is compiled once
per (variant, shard) so the fully unrolled kernels compile in parallel, and
``lup_persist_inst.hip`` once per (fused variant, size class).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "liblufqhe.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liblufqhe_oracle.so"

ARCH = "gfx950"
N_VARIANTS = 3
N_SHARDS = 8
N_PERSIST_VARIANTS = 2  # persistent mixed-size kernels: fused CROUT_REF and ABS
N_PERSIST_CLASSES = 5   # size classes of lup_persistent.hpp

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",  # every rounding step of the kernels is explicit
    "-mllvm", "-pragma-unroll-threshold=1000000",  # fully unroll the step loop up to n = 64
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _headers() -> list[Path]:
    return sorted(CSRC.glob("*.hpp")) + sorted((ROOT / "include").glob("*.h*"))


def _stale(obj: Path, deps: list[Path]) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}")


def _compile_jobs() -> list[tuple[Path, list[str], list[Path]]]:
    hdrs = _headers()
    # the kernel instantiation units include only the kernel headers
    kern_hdrs = sorted(CSRC.glob("*.hpp"))
    jobs = []
    # the persistent class kernels are the longest single compiles: first
    for v in range(N_PERSIST_VARIANTS):
        for c in range(N_PERSIST_CLASSES):
            obj = BUILD / f"lup_persist_v{v}_c{c}.o"
            defs = [f"-DLUFQHE_VARIANT={v}", f"-DLUFQHE_CLASS={c}"]
            jobs.append((obj, defs + ["-c", str(CSRC / "lup_persist_inst.hip")],
                         kern_hdrs + [CSRC / "lup_persist_inst.hip"]))
    for v in range(N_VARIANTS):
        for s in range(N_SHARDS):
            obj = BUILD / f"lup_inst_v{v}_s{s}.o"
            defs = [f"-DLUFQHE_VARIANT={v}", f"-DLUFQHE_SHARD={s}", f"-DLUFQHE_SHARDS={N_SHARDS}"]
            jobs.append((obj, defs + ["-c", str(CSRC / "lup_inst.hip")],
                         kern_hdrs + [CSRC / "lup_inst.hip"]))
    for src in sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp")):
        if src.name in ("lup_inst.hip", "lup_persist_inst.hip"):
            continue
        obj = BUILD / (src.stem + ".o")
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        jobs.append((obj, lang + ["-c", str(src)], hdrs + [src]))
    return jobs


def build_library(force: bool = True, jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    cc = hipcc()
    work = _compile_jobs()
    todo = [(o, a) for (o, a, deps) in work if force or _stale(o, deps)]
    n = jobs or min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=n) as ex:
        futs = [ex.submit(_run, [cc, *HIP_FLAGS, *args, "-o", str(obj)], verbose) for obj, args in todo]
        for f in futs:
            f.result()
    objs = [o for (o, _, _) in work]
    if force or todo or _stale(LIB, objs):
        tmp = LIB.with_suffix(".so.tmp")
        _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)], verbose)
        os.replace(tmp, LIB)
    return LIB


DRIVER_SRC = ROOT / "tests" / "cpp" / "lucore_gpu_driver.cpp"
DRIVER = ROOT / "tests" / "cpp" / "lucore_gpu_driver"


def build_driver(force: bool = True, verbose: bool = False) -> Path:
    """C++ test driver of include/lucore_gpu.hpp, linked against the library."""
    deps = [DRIVER_SRC, LIB, ROOT / "include" / "lucore_gpu.hpp", ROOT / "include" / "lufqhe.h"]
    if force or _stale(DRIVER, deps):
        _run([hipcc(), "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(DRIVER_SRC), "-o", str(DRIVER),
              f"-L{PKG}", "-llufqhe", "-Wl,-rpath,$ORIGIN/../../lion-codes_amd"], verbose)
    return DRIVER


CPU_DRIVER_SRC = ROOT / "tests" / "cpp" / "lucore_cpu_driver.cpp"
CPU_DRIVER = ROOT / "tests" / "cpp" / "lucore_cpu_driver"


def build_cpu_driver(force: bool = True, verbose: bool = False) -> Path:
    """C++ test driver of include/lucore_cpu.hpp, linked against the library."""
    deps = [CPU_DRIVER_SRC, LIB, ROOT / "include" / "lucore_cpu.hpp", ROOT / "include" / "lufqhe.h"]
    if force or _stale(CPU_DRIVER, deps):
        _run([hipcc(), "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(CPU_DRIVER_SRC), "-o",
              str(CPU_DRIVER), f"-L{PKG}", "-llufqhe", "-Wl,-rpath,$ORIGIN/../../lion-codes_amd"],
             verbose)
    return CPU_DRIVER


def build_oracle(force: bool = True, verbose: bool = False) -> Path:
    src = ORACLE_DIR / "lufqhe_oracle.c"
    deps = [src, ORACLE_DIR / "lufqhe_oracle.h"]
    if force or _stale(ORACLE_LIB, deps):
        tmp = ORACLE_LIB.with_suffix(".so.tmp")
        _run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
              "-o", str(tmp), str(src), "-lm"], verbose)
        os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build(force: bool = True, jobs: int | None = None, verbose: bool = False) -> None:
    build_oracle(force=force, verbose=verbose)
    build_library(force=force, jobs=jobs, verbose=verbose)
    build_driver(force=force, verbose=verbose)
    build_cpu_driver(force=force, verbose=verbose)


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose="-v" in sys.argv)
