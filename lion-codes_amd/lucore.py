"""Python host binding of the batched LUP-determinant engine.

Mirrors the reference's GPU entry point for this path -- a batch of small
complex matrices in, one determinant per matrix out (LUCoreGPU::CALL,
LU-fqhe/src/LUCoreGPU.cu:38-106) -- over the C ABI of ``include/lufqhe.h``.
PyTorch provides device memory and streams only; all arithmetic runs in the
HIP library.  There is no fallback: without the built library every call
raises.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from pathlib import Path

import torch  # imported before the library so both share one HIP runtime

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "liblufqhe.so"

POLICIES = {"crout_ref": 0, "abs": 1}
ARITHS = {"fused": 0, "ref": 1}

SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_UNSUPPORTED = 2
ERROR_HIP = 3
ERROR_NOT_BUILT = 4
ERROR_OUT_OF_BUDGET = 5

_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


class LufqheError(RuntimeError):
    """A non-success status from the C ABI."""

    def __init__(self, fn: str, status: int):
        lib = load()
        msg = lib.lufqhe_status_string(status).decode()
        hip = lib.lufqhe_last_hip_error().decode()
        super().__init__(f"{fn}: status {status} ({msg}){' / ' + hip if hip else ''}")
        self.status = status


def load() -> ctypes.CDLL:
    """Load liblufqhe.so (built in-tree by ``build()``); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.lufqhe_version.restype = _i32
    lib.lufqhe_status_string.restype = ctypes.c_char_p
    lib.lufqhe_status_string.argtypes = [_i32]
    lib.lufqhe_last_hip_error.restype = ctypes.c_char_p
    lib.lufqhe_det_batched.restype = _i32
    lib.lufqhe_det_batched.argtypes = [_vp, _i64, _i32, _i64, _i64, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _i32, _i32, _vp]
    _lib = lib
    return lib


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _stream_handle(stream: torch.cuda.Stream | None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


@dataclass
class DetResult:
    """Per-matrix determinants: det = mant * 2**exp."""

    mant: torch.Tensor      # complex128 [B]
    exp: torch.Tensor       # int32 [B]
    logabs: torch.Tensor    # float64 [B]
    phase: torch.Tensor     # float64 [B]
    info: torch.Tensor      # int32 [B]
    piv: torch.Tensor | None = None  # int8 [B, n]

    def det(self) -> torch.Tensor:
        """Plain complex determinant (overflows to inf where the reference's
        plain product would)."""
        return torch.ldexp(self.mant, self.exp.to(torch.float64))


def alloc_result(batch: int, n: int, device, pivots: bool = False) -> DetResult:
    return DetResult(
        mant=torch.empty(batch, dtype=torch.complex128, device=device),
        exp=torch.empty(batch, dtype=torch.int32, device=device),
        logabs=torch.empty(batch, dtype=torch.float64, device=device),
        phase=torch.empty(batch, dtype=torch.float64, device=device),
        info=torch.empty(batch, dtype=torch.int32, device=device),
        piv=torch.empty(batch, n, dtype=torch.int8, device=device) if pivots else None,
    )


def det_batched(A: torch.Tensor, policy: str = "crout_ref", arith: str = "fused",
                pivots: bool = False, out: DetResult | None = None,
                stream: torch.cuda.Stream | None = None) -> DetResult:
    """Determinants of a uniform batch ``A`` (complex128, shape [B, n, n], on
    the GPU, unit column stride; row stride = ld, matrix stride = stride)."""
    if A.dtype != torch.complex128 or not A.is_cuda or A.dim() != 3:
        raise ValueError("A must be a complex128 CUDA tensor of shape [B, n, n]")
    B, n, n2 = A.shape
    if n != n2 or A.stride(2) != 1:
        raise ValueError("A must hold square matrices with unit column stride")
    res = out if out is not None else alloc_result(B, n, A.device, pivots)
    st = load().lufqhe_det_batched(
        _ptr(A), B, n, A.stride(1), A.stride(0), _ptr(res.mant), _ptr(res.exp),
        _ptr(res.logabs), _ptr(res.phase), _ptr(res.info), _ptr(res.piv),
        POLICIES[policy], ARITHS[arith], _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_det_batched", st)
    return res


def logabs_phase_from(mant_re, mant_im, exp):
    """log|det| and arg det from (mant, exp) -- host helper for checks."""
    m = math.hypot(mant_re, mant_im)
    if m == 0.0:
        return -math.inf, 0.0
    return math.log(m) + exp * math.log(2.0), math.atan2(mant_im, mant_re)
