"""AddressSanitizer and UndefinedBehaviorSanitizer runs of the CPU code: the
oracle (tests/sanitize/oracle_driver.c over oracle/lufqhe_oracle.c) and the
library's host-only parts (tests/sanitize/host_driver.cpp over the CPU
Crout, RunInfo and the shape-table reader).  Built with the sanitizers on
the host side only; no GPU is involved."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    assert r.returncode == 0, f"{' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}"
    return r


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_driver"
    run(["gcc", *SAN, "-fopenmp", "-ffp-contract=off", str(ROOT / "oracle" / "lufqhe_oracle.c"),
         str(ROOT / "tests" / "sanitize" / "oracle_driver.c"), "-o", str(exe), "-lm"])
    r = run([str(exe)], env=dict(ENV, OMP_NUM_THREADS="3"))
    assert "oracle driver: ok" in r.stdout


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc (for the HIP headers of the host code) is not installed")
def test_host_code_under_sanitizers(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "host_driver"
    host_san = []
    for f in SAN:
        host_san += ["-Xarch_host", f] if f.startswith("-fsanitize") or f.startswith("-fno-sanitize") else [f]
    csrc = ROOT / "lion-codes_amd" / "csrc"
    run([hipcc, *host_san, "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}",
         str(csrc / "lucore_cpu.cpp"), str(csrc / "run_info.cpp"), str(csrc / "lufqhe_table.cpp"),
         str(ROOT / "tests" / "sanitize" / "host_driver.cpp"), "-o", str(exe), "-lpthread"])
    r = run([str(exe), str(ROOT / "lion-codes_amd" / "shapes_gfx950.json")],
            env=dict(ENV, TMPDIR=str(tmp_path)))
    assert "host driver: ok" in r.stdout
