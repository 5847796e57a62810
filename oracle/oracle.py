"""ctypes front end of the CPU oracle (liblufqhe_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, as the checker / baseline -- never by the
product.  Inputs are numpy complex arrays; see lufqhe_oracle.c for what each
restatement follows in the reference.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liblufqhe_oracle.so"
POLICY = {"crout_ref": 0, "abs": 1}
ARITH = {"fused": 0, "ref": 1}

_lib = None
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.oracle_max_threads.restype = _i32
        lib.oracle_crout_ref_batch_z.argtypes = [_vp, _i64, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _i32]
        lib.oracle_crout_ref_batch_c.argtypes = [_vp, _i64, _i32, _i32, _i64, _vp, _vp, _i32]
        lib.oracle_lup_mirror_batch_z.argtypes = [_vp, _i64, _i32, _i32, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _i32]
        lib.oracle_lup_ld_batch_z.argtypes = [_vp, _i64, _i32, _i32, _i64, _vp, _vp, _vp, _i32]
        lib.oracle_lup_mirror_mixed_z.argtypes = [_vp, _i64, _vp, _vp, _i32, _i32, _vp, _vp, _vp, _i32]
        lib.oracle_lup_ld_mixed_z.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _i32]
        lib.oracle_crout_ref_pairsum_z.argtypes = [_vp, _i64, _i32, _i32, _vp]
        lib.oracle_crout_ref_gap_batch_z.argtypes = [_vp, _i64, _i32, _i32, _i64, _vp, _vp, _vp,
                                                     _i32]
        lib.oracle_quot_rn.restype = ctypes.c_double
        lib.oracle_quot_rn.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _batch(A, dtype):
    A = np.ascontiguousarray(A, dtype=dtype)
    if A.ndim == 2:
        A = A[None]
    count, n, n2 = A.shape
    assert n == n2
    return A, count, n


def crout_ref(A, nthreads: int = 0, pivots: bool = False):
    """Restated LUCoreCPU::determinant at complex128: dict(det, mant, exp, info, piv)."""
    A, count, n = _batch(A, np.complex128)
    det = np.empty(count, np.complex128)
    mant = np.empty(count, np.complex128)
    exp = np.empty(count, np.int32)
    info = np.empty(count, np.int32)
    piv = np.empty((count, n), np.int32) if pivots else None
    load().oracle_crout_ref_batch_z(_p(A), count, n, n, n * n, _p(det), _p(mant), _p(exp),
                                    _p(info), _p(piv), nthreads)
    return dict(det=det, mant=mant, exp=exp, info=info, piv=piv)


def crout_ref_gaps(A, nthreads: int = 0):
    """Pivot sequence of the restated Crout and, per step, the relative margin
    (c_best - c_second) / c_best of its choice: dict(piv, gap, info), piv and
    gap of shape (count, n); gap is 0 at an exact tie, inf at the last step."""
    A, count, n = _batch(A, np.complex128)
    piv = np.empty((count, n), np.int32)
    gap = np.empty((count, n), np.float64)
    info = np.empty(count, np.int32)
    load().oracle_crout_ref_gap_batch_z(_p(A), count, n, n, n * n, _p(piv), _p(gap), _p(info),
                                        nthreads)
    return dict(piv=piv, gap=gap, info=info)


def pivot_parity(piv_gpu, ref: dict, margin: float) -> dict:
    """Compares a pivot table with crout_ref_gaps' output.  A matrix whose
    sequences differ is an excluded near-tie when the reference's margin at
    the first differing step is below `margin`, else a mismatch."""
    piv_gpu = np.asarray(piv_gpu, np.int32)
    ok = ref["info"] == 0
    diff = (piv_gpu != ref["piv"]) & ok[:, None]
    rows = np.nonzero(diff.any(axis=1))[0]
    first = diff[rows].argmax(axis=1)
    near = ref["gap"][rows, first] < margin
    ties = int((ref["gap"][ok] == 0).any(axis=1).sum())
    return {"matrices": int(len(piv_gpu)), "compared": int(ok.sum()),
            "differing": int(len(rows)), "excluded_near_ties": int(near.sum()),
            "mismatches": int((~near).sum()), "margin": margin,
            "matrices_with_exact_ties": ties,
            "mismatch_index": [int(i) for i in rows[~near][:10]]}


def quot_rn(x: float, s: float) -> float:
    """The kernel's fused CROUT_REF criterion x / s (corrected product)."""
    y = 1.0 / s if s <= 2.0 ** 960 else float("nan")
    return load().oracle_quot_rn(x, s, y)


def crout_ref_c64(A, nthreads: int = 0):
    """Restated LUCoreCPU::determinant at the reference's own complex64."""
    A, count, n = _batch(A, np.complex64)
    det = np.empty(count, np.complex64)
    info = np.empty(count, np.int32)
    load().oracle_crout_ref_batch_c(_p(A), count, n, n, n * n, _p(det), _p(info), nthreads)
    return dict(det=det, info=info)


def mirror(A, policy: str = "crout_ref", arith: str = "fused", nthreads: int = 0,
           pivots: bool = True):
    """Right-looking LUP in the kernel's operation order: dict(mant, exp, info, piv)."""
    A, count, n = _batch(A, np.complex128)
    mant = np.empty(count, np.complex128)
    exp = np.empty(count, np.int32)
    info = np.empty(count, np.int32)
    piv = np.empty((count, n), np.int32) if pivots else None
    load().oracle_lup_mirror_batch_z(_p(A), count, n, n, n * n, POLICY[policy], ARITH[arith],
                                     _p(mant), _p(exp), _p(info), _p(piv), nthreads)
    return dict(mant=mant, exp=exp, info=info, piv=piv)


def truth(A, nthreads: int = 0):
    """Long-double partial-pivoting log|det| and arg det: dict(logabs, phase, info)."""
    A, count, n = _batch(A, np.complex128)
    la = np.empty(count)
    ph = np.empty(count)
    info = np.empty(count, np.int32)
    load().oracle_lup_ld_batch_z(_p(A), count, n, n, n * n, _p(la), _p(ph), _p(info), nthreads)
    return dict(logabs=la, phase=ph, info=info)


def mirror_mixed(flat, ns, offs, policy="crout_ref", arith="fused", nthreads: int = 0):
    flat = np.ascontiguousarray(flat, np.complex128)
    ns = np.ascontiguousarray(ns, np.int32)
    offs = np.ascontiguousarray(offs, np.int64)
    count = len(ns)
    mant = np.empty(count, np.complex128)
    exp = np.empty(count, np.int32)
    info = np.empty(count, np.int32)
    load().oracle_lup_mirror_mixed_z(_p(flat), count, _p(ns), _p(offs), POLICY[policy],
                                     ARITH[arith], _p(mant), _p(exp), _p(info), nthreads)
    return dict(mant=mant, exp=exp, info=info)


def truth_mixed(flat, ns, offs, nthreads: int = 0):
    flat = np.ascontiguousarray(flat, np.complex128)
    ns = np.ascontiguousarray(ns, np.int32)
    offs = np.ascontiguousarray(offs, np.int64)
    count = len(ns)
    la = np.empty(count)
    ph = np.empty(count)
    info = np.empty(count, np.int32)
    load().oracle_lup_ld_mixed_z(_p(flat), count, _p(ns), _p(offs), _p(la), _p(ph), _p(info),
                                 nthreads)
    return dict(logabs=la, phase=ph, info=info)


def crout_ref_pairsum(A):
    """LUCoreCPU::CALL analogue: sum_i det(M_2i) * det(M_2i+1)."""
    A, count, n = _batch(A, np.complex128)
    out = np.empty(1, np.complex128)
    load().oracle_crout_ref_pairsum_z(_p(A), count, n, n, _p(out))
    return out[0]


def max_threads() -> int:
    return load().oracle_max_threads()


def log_form(mant, exp):
    """log|det| and arg det from (mant, exp) arrays."""
    mant = np.asarray(mant)
    with np.errstate(divide="ignore"):
        la = np.log(np.abs(mant)) + np.asarray(exp, np.float64) * np.log(2.0)
    ph = np.angle(mant)
    ph = np.where(mant == 0, 0.0, ph)
    return la, ph
