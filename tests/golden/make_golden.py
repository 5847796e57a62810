"""Writes the golden fixtures under tests/golden/ (run once, in the dev container).

    python tests/golden/make_golden.py

Every expected value is an mpmath determinant at 60 significant digits of the
exact float64 inputs stored beside it, so the fixtures pin the oracle without
relying on any of the code under test:

  gauss.npz          N(0,1) complex matrices, n in {1,2,3,4,5,7,8,12,16,24,32}
  near_singular.npz  n in {8,16,32}: one row = alpha*row_a + beta*row_b + 1e-8*noise
  fqhe.npz           FQHE matrices (z_i^j exp(-|z_i|^2/4)), n in {10,16,24,32,40},
                     also checked against the analytic Vandermonde form
  ref_fixture16.npz  the first 32 matrices of the reference's own input file
                     LU-fqhe/inc/1000_16x16.bin (complex64, entries in (0,1))

Fields: A (complex128, [count, n, n]) per size key "A_<n>", and "logabs_<n>",
"phase_<n>", "det_<n>" (complex128; inf if it overflows float64).
"""
from __future__ import annotations

import sys
from pathlib import Path

import mpmath as mp
import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
mp.mp.dps = 60

REF_FIXTURE = Path("/root/reference/LU-fqhe/inc/1000_16x16.bin")


def mp_det(A: np.ndarray):
    n = A.shape[0]
    M = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            M[i, j] = mp.mpc(mp.mpf(float(A[i, j].real)), mp.mpf(float(A[i, j].imag)))
    d = mp.det(M)
    la = mp.log(abs(d))
    ph = mp.arg(d)
    try:
        dc = complex(d)
    except OverflowError:
        dc = complex(np.inf, np.inf)
    return float(la), float(ph), dc


def pack(out: dict, key: str, mats: list[np.ndarray]) -> None:
    A = np.stack(mats).astype(np.complex128)
    la, ph, dd = zip(*(mp_det(a) for a in A))
    out[f"A_{key}"] = A
    out[f"logabs_{key}"] = np.array(la)
    out[f"phase_{key}"] = np.array(ph)
    out[f"det_{key}"] = np.array(dd, np.complex128)


def gauss() -> dict:
    rng = np.random.default_rng(20240601)
    out = {}
    for n in (1, 2, 3, 4, 5, 7, 8, 12, 16, 24, 32):
        cnt = 6 if n <= 16 else 3
        mats = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(cnt)]
        pack(out, str(n), mats)
    return out


def near_singular() -> dict:
    rng = np.random.default_rng(20240602)
    out = {}
    for n in (8, 16, 32):
        mats = []
        for _ in range(3):
            M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
            r, a, b = rng.choice(n, 3, replace=False)
            al = complex(*rng.standard_normal(2))
            be = complex(*rng.standard_normal(2))
            M[r] = al * M[a] + be * M[b] + 1e-8 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
            mats.append(M)
        pack(out, str(n), mats)
    return out


def fqhe() -> dict:
    import importlib
    wl = importlib.import_module("lion-codes_amd.workloads")
    rng = np.random.default_rng(20240603)
    out = {}
    for n in (10, 16, 24, 32, 40):
        zs = [wl.fqhe_positions(n, rng) for _ in range(2)]
        pack(out, str(n), [wl.fqhe_matrix(z) for z in zs])
        out[f"z_{n}"] = np.stack(zs)
    return out


def ref_fixture() -> dict:
    raw = np.fromfile(REF_FIXTURE, dtype=np.float32)
    A = raw.view(np.complex64).reshape(1000, 16, 16)[:32]
    out = {}
    pack(out, "16", list(A.astype(np.complex128)))
    out["A32_16"] = A  # the original complex64 values
    return out


def main() -> None:
    for name, fn in (("gauss", gauss), ("near_singular", near_singular), ("fqhe", fqhe),
                     ("ref_fixture16", ref_fixture)):
        d = fn()
        np.savez_compressed(HERE / f"{name}.npz", **d)
        print(name, {k: v.shape for k, v in d.items() if k.startswith("A")})


if __name__ == "__main__":
    main()
