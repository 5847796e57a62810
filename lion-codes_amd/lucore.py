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
    lib.lufqhe_det_inplace_workspace_bytes.restype = _i64
    lib.lufqhe_det_inplace_workspace_bytes.argtypes = [_i64]
    lib.lufqhe_det_inplace.restype = _i32
    lib.lufqhe_det_inplace.argtypes = [_vp, _i64, _i32, _i64, _i64, _i32, _i32, _vp, _vp]
    lib.lufqhe_pair_sum_workspace_bytes.restype = _i64
    lib.lufqhe_pair_sum_workspace_bytes.argtypes = []
    lib.lufqhe_global_reduction.restype = _i32
    lib.lufqhe_global_reduction.argtypes = [_vp, _i64, _i64, _vp, _vp, _vp]
    lib.lufqhe_pair_sum.restype = _i32
    lib.lufqhe_pair_sum.argtypes = [_vp, _vp, _i64, _vp, _vp, _vp, _vp]
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


def det_inplace(A: torch.Tensor, pair_product: bool = True, policy: str = "crout_ref",
                workspace: torch.Tensor | None = None,
                stream: torch.cuda.Stream | None = None) -> None:
    """The reference's in-place contract (LUCoreGPU::CALL with WRITE_DET and
    PERFORM_PROD): A[m,0,0] = det(A[m]), A[2i,0,0] = det(A[2i]) * det(A[2i+1])."""
    if A.dtype != torch.complex128 or not A.is_cuda or A.dim() != 3 or A.stride(2) != 1:
        raise ValueError("A must be a complex128 CUDA tensor [B, n, n] with unit column stride")
    B, n, _ = A.shape
    st = load().lufqhe_det_inplace(_ptr(A), B, n, A.stride(1), A.stride(0), int(pair_product),
                                   POLICIES[policy], _ptr(workspace), _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_det_inplace", st)


def global_reduction(A: torch.Tensor, workspace: torch.Tensor | None = None,
                     stream: torch.cuda.Stream | None = None) -> torch.Tensor:
    """sum_i A[2i,0,0] as one complex128 on the device (globalReduction)."""
    out = torch.empty(1, dtype=torch.complex128, device=A.device)
    st = load().lufqhe_global_reduction(_ptr(A), A.shape[0], A.stride(0), _ptr(workspace),
                                        _ptr(out), _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_global_reduction", st)
    return out


def pair_sum(res: DetResult, workspace: torch.Tensor | None = None,
             stream: torch.cuda.Stream | None = None) -> tuple[torch.Tensor, torch.Tensor]:
    """sum_i det_2i * det_2i+1 = mant * 2**exp, from a DetResult (overflow free)."""
    mant = torch.empty(1, dtype=torch.complex128, device=res.mant.device)
    exp = torch.empty(1, dtype=torch.int32, device=res.mant.device)
    st = load().lufqhe_pair_sum(_ptr(res.mant), _ptr(res.exp), res.mant.shape[0],
                                _ptr(workspace), _ptr(mant), _ptr(exp), _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_pair_sum", st)
    return mant, exp


def logabs_phase_from(mant_re, mant_im, exp):
    """log|det| and arg det from (mant, exp) -- host helper for checks."""
    m = math.hypot(mant_re, mant_im)
    if m == 0.0:
        return -math.inf, 0.0
    return math.log(m) + exp * math.log(2.0), math.atan2(mant_im, mant_re)
