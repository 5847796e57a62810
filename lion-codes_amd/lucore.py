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
# Launch shapes chosen by the autotuner on the target device (see autotune.py).
SHAPE_TABLE = _PKG / "shapes_gfx950.json"

POLICIES = {"crout_ref": 0, "abs": 1}
ARITHS = {"fused": 0, "ref": 1}

SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_UNSUPPORTED = 2
ERROR_HIP = 3
ERROR_NOT_BUILT = 4
ERROR_OUT_OF_BUDGET = 5
ERROR_IO = 6
ERROR_TOO_LARGE = 7
FILE_FORMATS = {"complex64": 0, "complex128": 1}

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
    lib.lufqhe_det_inplace_static.restype = _i32
    lib.lufqhe_det_inplace_static.argtypes = [_vp, _i64, _i32, _i64, _i64, _i32, _i32, _vp]
    lib.lufqhe_pair_sum_workspace_bytes.restype = _i64
    lib.lufqhe_pair_sum_workspace_bytes.argtypes = []
    lib.lufqhe_global_reduction.restype = _i32
    lib.lufqhe_global_reduction.argtypes = [_vp, _i64, _i64, _vp, _vp, _vp]
    lib.lufqhe_pair_sum.restype = _i32
    lib.lufqhe_pair_sum.argtypes = [_vp, _vp, _i64, _vp, _vp, _vp, _vp]
    lib.lufqhe_shape_count.restype = _i32
    lib.lufqhe_shape_count.argtypes = [_i32]
    lib.lufqhe_shape_info.restype = _i32
    lib.lufqhe_shape_info.argtypes = [_i32, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                      ctypes.POINTER(_i32)]
    lib.lufqhe_set_shape.restype = _i32
    lib.lufqhe_set_shape.argtypes = [_i32, _i32, _i32, _i32]
    lib.lufqhe_get_shape.restype = _i32
    lib.lufqhe_get_shape.argtypes = [_i32, _i32, _i32]
    lib.lufqhe_mixed_plan_create.restype = _i32
    lib.lufqhe_mixed_plan_create.argtypes = [_vp, _vp, _i64, _i32, _i32, ctypes.POINTER(_vp)]
    lib.lufqhe_mixed_plan_destroy.restype = _i32
    lib.lufqhe_mixed_plan_destroy.argtypes = [_vp]
    lib.lufqhe_mixed_plan_buckets.restype = _i32
    lib.lufqhe_mixed_plan_buckets.argtypes = [_vp, _vp, _vp, _vp, _i32]
    lib.lufqhe_mixed_execute.restype = _i32
    lib.lufqhe_mixed_execute.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _vp]
    lib.lufqhe_fqhe_det.restype = _i32
    lib.lufqhe_fqhe_det.argtypes = [_vp, _i64, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _i32, _i32, _vp]
    lib.lufqhe_server_create.restype = _i32
    lib.lufqhe_server_create.argtypes = [_i32, _vp, _i32, ctypes.c_uint64, _i32,
                                         ctypes.POINTER(_vp)]
    lib.lufqhe_server_destroy.argtypes = [_vp]
    lib.lufqhe_server_reset.restype = _i32
    lib.lufqhe_server_reset.argtypes = [_vp]
    lib.lufqhe_server_walker.restype = _i32
    lib.lufqhe_server_walker.argtypes = [_vp, _i32, ctypes.POINTER(_i32), _vp]
    lib.lufqhe_server_perms.restype = _i32
    lib.lufqhe_server_perms.argtypes = [_vp, _i32, _vp]
    lib.lufqhe_server_run.restype = _i32
    lib.lufqhe_server_run.argtypes = [_vp, ctypes.POINTER(ServeConfig), ctypes.POINTER(ServeStats),
                                      _vp]
    lib.lufqhe_file_chunk.restype = _i32
    lib.lufqhe_file_chunk.argtypes = [_i64, _i32, _i32, _i32, ctypes.POINTER(_i64),
                                      ctypes.POINTER(_i64)]
    lib.lufqhe_read_matrices.restype = _i32
    lib.lufqhe_read_matrices.argtypes = [ctypes.c_char_p, _i32, _i64, _i32, _i32, _i32, _vp, _i64,
                                         ctypes.POINTER(_i64), _vp]
    lib.lufqhe_load_shape_table.restype = _i32
    lib.lufqhe_load_shape_table.argtypes = [ctypes.c_char_p]
    lib.lufqhe_cpu_crout_det.restype = _i32
    lib.lufqhe_cpu_crout_det.argtypes = [_vp, _i64, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp,
                                         _i32]
    lib.lufqhe_cpu_pair_sum.restype = _i32
    lib.lufqhe_cpu_pair_sum.argtypes = [_vp, _i64, _i32, _i64, _i64, _vp, _vp, _i32]
    lib.lufqhe_mc_split.restype = _i32
    lib.lufqhe_mc_split.argtypes = [_i64, _i32, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
    lib.lufqhe_mc_lu_run.restype = _i32
    lib.lufqhe_mc_lu_run.argtypes = [_vp, _i64, _i32, _i32, _i32, _i32, _vp,
                                     ctypes.POINTER(McStats)]
    lib.lufqhe_fqhe_build.restype = _i32
    lib.lufqhe_fqhe_build.argtypes = [_vp, _i64, _i32, _i64, _vp, _vp, _vp]
    _lib = lib
    apply_shape_table()
    return lib


def apply_shape_table(path: Path | None = None) -> int:
    """Select the tuned launch shape of every (policy, n) listed in the table
    written by autotune.py; returns the number of entries applied."""
    p = Path(path) if path is not None else SHAPE_TABLE
    if not p.exists():
        return 0
    import json
    table = json.loads(p.read_text())
    applied = 0
    for e in table.get("entries", []):
        n, sh = int(e["n"]), int(e["shape"])
        # an entry applies only to the candidate it was measured on
        if not 0 <= sh < shape_count(n):
            continue
        if shape_info(n, sh) != (e["lanes"], e["lane_cols"], e.get("waves", 1)):
            continue
        if set_shape(e["policy"], "fused", n, sh):
            applied += 1
    return applied


def shape_count(n: int) -> int:
    return load().lufqhe_shape_count(n)


def shape_info(n: int, shape: int) -> tuple[int, int, int]:
    """(lanes per matrix, lane-columns, waves per SIMD the register
    allocation allows) of candidate `shape` for size n."""
    lanes, cols, waves = _i32(), _i32(), _i32()
    st = load().lufqhe_shape_info(n, shape, ctypes.byref(lanes), ctypes.byref(cols),
                                ctypes.byref(waves))
    if st != SUCCESS:
        raise LufqheError("lufqhe_shape_info", st)
    return lanes.value, cols.value, waves.value


def set_shape(policy: str, arith: str, n: int, shape: int) -> bool:
    return load().lufqhe_set_shape(POLICIES[policy], ARITHS[arith], n, shape) == SUCCESS


def get_shape(policy: str, arith: str, n: int) -> int:
    return load().lufqhe_get_shape(POLICIES[policy], ARITHS[arith], n)


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _check_device_tensor(name: str, t: torch.Tensor, dtype, numel: int, device) -> None:
    """A kernel writes or reads `numel` elements of `t`: refuse anything that
    would put that access out of bounds or on another device."""
    if t.dtype != dtype or not t.is_cuda or t.device != device:
        raise ValueError(f"{name} must be a {dtype} GPU tensor on {device}")
    if not t.is_contiguous() or t.numel() < numel:
        raise ValueError(f"{name} must be contiguous with at least {numel} elements")


def _check_workspace(ws: torch.Tensor | None, nbytes: int, device) -> None:
    if ws is None:
        return
    if not ws.is_cuda or ws.device != device or not ws.is_contiguous():
        raise ValueError(f"workspace must be a contiguous tensor on {device}")
    if ws.numel() * ws.element_size() < nbytes:
        raise ValueError(f"workspace must hold at least {nbytes} bytes")


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

    def check(self, batch: int, n: int, device) -> None:
        """Raise unless every output holds `batch` results on `device` (mant is
        required; the library skips an output given as None)."""
        if self.mant is None:
            raise ValueError("mant is required")
        fields = (("mant", torch.complex128, 1), ("exp", torch.int32, 1),
                  ("logabs", torch.float64, 1), ("phase", torch.float64, 1),
                  ("info", torch.int32, 1), ("piv", torch.int8, n))
        for name, dtype, per in fields:
            t = getattr(self, name)
            if t is not None:
                _check_device_tensor(name, t, dtype, batch * per, device)


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
    res.check(B, n, A.device)
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
    if (A.dtype != torch.complex128 or not A.is_cuda or A.dim() != 3 or A.stride(2) != 1
            or A.shape[1] != A.shape[2]):
        raise ValueError("A must be a complex128 CUDA tensor [B, n, n] with unit column stride")
    B, n, _ = A.shape
    _check_workspace(workspace, load().lufqhe_det_inplace_workspace_bytes(B), A.device)
    st = load().lufqhe_det_inplace(_ptr(A), B, n, A.stride(1), A.stride(0), int(pair_product),
                                   POLICIES[policy], _ptr(workspace), _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_det_inplace", st)


def global_reduction(A: torch.Tensor, workspace: torch.Tensor | None = None,
                     stream: torch.cuda.Stream | None = None) -> torch.Tensor:
    """sum_i A[2i,0,0] as one complex128 on the device (globalReduction)."""
    if A.dtype != torch.complex128 or not A.is_cuda or A.dim() != 3:
        raise ValueError("A must be a complex128 CUDA tensor of shape [B, n, n]")
    _check_workspace(workspace, load().lufqhe_pair_sum_workspace_bytes(), A.device)
    out = torch.empty(1, dtype=torch.complex128, device=A.device)
    st = load().lufqhe_global_reduction(_ptr(A), A.shape[0], A.stride(0), _ptr(workspace),
                                        _ptr(out), _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_global_reduction", st)
    return out


def pair_sum(res: DetResult, workspace: torch.Tensor | None = None,
             stream: torch.cuda.Stream | None = None) -> tuple[torch.Tensor, torch.Tensor]:
    """sum_i det_2i * det_2i+1 = mant * 2**exp, from a DetResult (overflow free)."""
    dev = res.mant.device
    _check_device_tensor("mant", res.mant, torch.complex128, res.mant.numel(), dev)
    _check_device_tensor("exp", res.exp, torch.int32, res.mant.numel(), dev)
    _check_workspace(workspace, load().lufqhe_pair_sum_workspace_bytes(), dev)
    mant = torch.empty(1, dtype=torch.complex128, device=res.mant.device)
    exp = torch.empty(1, dtype=torch.int32, device=res.mant.device)
    st = load().lufqhe_pair_sum(_ptr(res.mant), _ptr(res.exp), res.mant.shape[0],
                                _ptr(workspace), _ptr(mant), _ptr(exp), _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_pair_sum", st)
    return mant, exp


def _check_fqhe_inputs(z: torch.Tensor, row_perm: torch.Tensor | None):
    if z.dtype != torch.complex128 or not z.is_cuda or z.dim() != 2 or z.stride(1) != 1:
        raise ValueError("z must be a complex128 CUDA tensor [count or 1, n], unit stride")
    if row_perm is not None and (row_perm.dtype != torch.uint8 or not row_perm.is_cuda
                                 or not row_perm.is_contiguous()):
        raise ValueError("row_perm must be a contiguous uint8 CUDA tensor [count, n]")


def fqhe_det(z: torch.Tensor, row_perm: torch.Tensor | None = None, count: int | None = None,
             policy: str = "crout_ref", arith: str = "fused", pivots: bool = False,
             out: DetResult | None = None,
             stream: torch.cuda.Stream | None = None) -> DetResult:
    """Determinants of FQHE matrices A_ij = z_i^j exp(-|z_i|^2/4) built on the
    device from positions ``z`` ([count, n] complex128, or [1, n] shared by
    all ``count`` matrices), rows optionally permuted by ``row_perm`` [count, n]."""
    _check_fqhe_inputs(z, row_perm)
    rows, n = z.shape
    z_stride = z.stride(0) if rows > 1 else 0
    count = rows if count is None else count
    if rows > 1 and count != rows:
        raise ValueError("count must equal z.shape[0] unless z holds one configuration")
    if row_perm is not None and tuple(row_perm.shape) != (count, n):
        raise ValueError("row_perm must have shape [count, n]")
    if row_perm is not None and row_perm.device != z.device:
        raise ValueError("row_perm must be on the device of z")
    res = out if out is not None else alloc_result(count, n, z.device, pivots)
    res.check(count, n, z.device)
    st = load().lufqhe_fqhe_det(_ptr(z), count, n, z_stride, _ptr(row_perm), _ptr(res.mant),
                                _ptr(res.exp), _ptr(res.logabs), _ptr(res.phase),
                                _ptr(res.info), _ptr(res.piv), POLICIES[policy], ARITHS[arith],
                                _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_fqhe_det", st)
    return res


def fqhe_build(z: torch.Tensor, row_perm: torch.Tensor | None = None, count: int | None = None,
               stream: torch.cuda.Stream | None = None) -> torch.Tensor:
    """The FQHE matrices fqhe_det builds, written out ([count, n, n] complex128)."""
    _check_fqhe_inputs(z, row_perm)
    rows, n = z.shape
    z_stride = z.stride(0) if rows > 1 else 0
    count = rows if count is None else count
    if rows > 1 and count != rows:
        raise ValueError("count must equal z.shape[0] unless z holds one configuration")
    if row_perm is not None and tuple(row_perm.shape) != (count, n):
        raise ValueError("row_perm must have shape [count, n]")
    if row_perm is not None and row_perm.device != z.device:
        raise ValueError("row_perm must be on the device of z")
    A = torch.empty(count, n, n, dtype=torch.complex128, device=z.device)
    st = load().lufqhe_fqhe_build(_ptr(z), count, n, z_stride, _ptr(row_perm), _ptr(A),
                                  _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_fqhe_build", st)
    return A


def cpu_crout_det(A, nthreads: int = 0, pivots: bool = False) -> dict:
    """The reference CPU Crout (LUCoreCPU::determinant) on the host through
    the C ABI: ``A`` a numpy complex128 array [B, n, n] (any row stride with
    unit column stride).  Returns dict(det, mant, exp, info, piv)."""
    import numpy as np
    A = np.asarray(A)
    if A.dtype != np.complex128 or A.ndim != 3 or A.strides[2] != 16:
        raise ValueError("A must be a complex128 numpy array [B, n, n] with unit column stride")
    B, n, n2 = A.shape
    if n != n2 or A.strides[1] % 16 or A.strides[0] % 16:
        raise ValueError("A must hold square matrices on complex128 strides")
    det = np.empty(B, np.complex128)
    mant = np.empty(B, np.complex128)
    exp = np.empty(B, np.int32)
    info = np.empty(B, np.int32)
    piv = np.empty((B, n), np.int8) if pivots else None
    st = load().lufqhe_cpu_crout_det(A.ctypes.data, B, n, A.strides[1] // 16, A.strides[0] // 16,
                                     det.ctypes.data, mant.ctypes.data, exp.ctypes.data,
                                     info.ctypes.data, None if piv is None else piv.ctypes.data,
                                     nthreads)
    if st != SUCCESS:
        raise LufqheError("lufqhe_cpu_crout_det", st)
    return dict(det=det, mant=mant, exp=exp, info=info, piv=piv)


def cpu_pair_sum(A, nthreads: int = 0) -> tuple[complex, int]:
    """LUCoreCPU::CALL's value on the host: sum_i det(A[2i]) * det(A[2i+1])
    (pair products added in matrix order) and the first pair holding a
    rejected row (-1 if none)."""
    import numpy as np
    A = np.ascontiguousarray(A, np.complex128)
    if A.ndim != 3 or A.shape[1] != A.shape[2]:
        raise ValueError("A must be a complex128 array [B, n, n]")
    B, n, _ = A.shape
    out = np.empty(1, np.complex128)
    bad = _i32()
    st = load().lufqhe_cpu_pair_sum(A.ctypes.data, B, n, n, n * n, out.ctypes.data,
                                    ctypes.byref(bad), nthreads)
    if st != SUCCESS:
        raise LufqheError("lufqhe_cpu_pair_sum", st)
    return complex(out[0]), bad.value


class McStats(ctypes.Structure):
    """lufqhe_mc_stats."""
    _fields_ = [("gpu_matrices", _i64), ("cpu_matrices", _i64), ("wall_s", ctypes.c_double),
                ("gpu_s", ctypes.c_double), ("cpu_s", ctypes.c_double)]


def mc_split(nelem: int, cpu_count: int) -> tuple[int, int]:
    """(GPU matrices, host matrices) of the reference's work split."""
    g, c = _i64(), _i64()
    st = load().lufqhe_mc_split(nelem, cpu_count, ctypes.byref(g), ctypes.byref(c))
    if st != SUCCESS:
        raise LufqheError("lufqhe_mc_split", st)
    return g.value, c.value


def mc_lu_run(A, cpu_count: int = 16, iterations: int = 1, policy: str = "crout_ref"):
    """The reference's Monte Carlo loop in LU mode on one GPU (lufqhe_mc_lu_run):
    ``A`` a numpy complex128 array [nelem, n, n] on the host.  Returns
    (results complex128 [iterations], stats dict)."""
    import numpy as np
    A = np.ascontiguousarray(A, np.complex128)
    if A.ndim != 3 or A.shape[1] != A.shape[2]:
        raise ValueError("A must be a complex128 array [nelem, n, n]")
    nelem, n, _ = A.shape
    out = np.empty(iterations, np.complex128)
    stats = McStats()
    st = load().lufqhe_mc_lu_run(A.ctypes.data, nelem, n, cpu_count, iterations,
                                 POLICIES[policy], out.ctypes.data, ctypes.byref(stats))
    if st != SUCCESS:
        raise LufqheError("lufqhe_mc_lu_run", st)
    return out, {name: getattr(stats, name) for name, _ in McStats._fields_}


class ServeConfig(ctypes.Structure):
    """lufqhe_serve_config."""
    _fields_ = [("rate", ctypes.c_double), ("requests", _i64), ("k_min", _i32), ("k_max", _i32),
                ("seed", ctypes.c_uint64), ("max_batch_requests", _i32), ("max_inflight", _i32),
                ("max_batch_dets", _i64), ("hbm_budget", _i64)]


class ServeStats(ctypes.Structure):
    """lufqhe_serve_stats."""
    _fields_ = [("requests", _i64), ("dets", _i64), ("batches", _i64), ("wall_s", ctypes.c_double),
                ("dets_per_s", ctypes.c_double), ("requests_per_s", ctypes.c_double),
                ("offered_rate", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("max_s", ctypes.c_double),
                ("mean_s", ctypes.c_double), ("mean_batch_dets", ctypes.c_double),
                ("budget_waits", _i64), ("peak_inflight_bytes", _i64), ("resident_bytes", _i64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# numpy view of lufqhe_serve_record
SERVE_RECORD_DTYPE = [("walker", "<i4"), ("particle", "<i4"), ("k", "<i4"), ("perm_start", "<i4"),
                      ("accepted", "<i4"), ("sum_exp", "<i4"), ("delta", "<c16"),
                      ("sum_mant", "<c16"), ("latency_s", "<f8")]


class Server:
    """Determinant server for Monte Carlo walkers (lufqhe_server_*): walker
    configurations and permutation tables live on the device; run() serves a
    Poisson stream of Monte Carlo steps with continuous batching."""

    def __init__(self, walker_n, perms_per_walker: int = 2000, seed: int = 0,
                 policy: str = "crout_ref"):
        import numpy as np
        self.walker_n = np.ascontiguousarray(walker_n, dtype=np.int32)
        self.walkers = len(self.walker_n)
        self.perms_per_walker = perms_per_walker
        self._h = _vp()
        st = load().lufqhe_server_create(self.walkers, self.walker_n.ctypes.data, perms_per_walker,
                                         seed, POLICIES[policy], ctypes.byref(self._h))
        if st != SUCCESS:
            raise LufqheError("lufqhe_server_create", st)

    def run(self, requests: int, rate: float = 0.0, k_min: int = 1, k_max: int = 2000,
            seed: int = 1, max_batch_requests: int = 256, max_inflight: int = 4,
            max_batch_dets: int = 1 << 20, hbm_budget: int = 200 << 30,
            records: bool = True):
        """Serve `requests` Monte Carlo steps; returns (stats dict, records
        structured array or None)."""
        import numpy as np
        cfg = ServeConfig(rate, requests, k_min, k_max, seed, max_batch_requests, max_inflight,
                          max_batch_dets, hbm_budget)
        stats = ServeStats()
        rec = np.zeros(requests, dtype=SERVE_RECORD_DTYPE) if records else None
        st = _lib.lufqhe_server_run(self._h, ctypes.byref(cfg), ctypes.byref(stats),
                                    None if rec is None else rec.ctypes.data)
        if st != SUCCESS:
            raise LufqheError("lufqhe_server_run", st)
        return stats.as_dict(), rec

    def reset(self) -> None:
        st = _lib.lufqhe_server_reset(self._h)
        if st != SUCCESS:
            raise LufqheError("lufqhe_server_reset", st)

    def walker(self, w: int):
        """Walker w's current configuration (complex128 [n])."""
        import numpy as np
        n = _i32()
        z = np.zeros(64, np.complex128)
        st = _lib.lufqhe_server_walker(self._h, w, ctypes.byref(n), z.ctypes.data)
        if st != SUCCESS:
            raise LufqheError("lufqhe_server_walker", st)
        return z[:n.value].copy()

    def perms(self, w: int):
        """Walker w's permutation table (uint8 [perms_per_walker, n])."""
        import numpy as np
        t = np.zeros((self.perms_per_walker, 64), np.uint8)
        st = _lib.lufqhe_server_perms(self._h, w, t.ctypes.data)
        if st != SUCCESS:
            raise LufqheError("lufqhe_server_perms", st)
        return t[:, :int(self.walker_n[w])].copy()

    def close(self) -> None:
        if self._h:
            _lib.lufqhe_server_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def file_chunk(count: int, n: int, rank: int = 0, nodes: int = 1) -> tuple[int, int]:
    """(first scalar, number of scalars) rank `rank` of `nodes` reads from the
    reference's matrix file (LU-fqhe/src/PermHandler.cpp:97-100)."""
    first, length = _i64(), _i64()
    st = load().lufqhe_file_chunk(count, n, rank, nodes, ctypes.byref(first), ctypes.byref(length))
    if st != SUCCESS:
        raise LufqheError("lufqhe_file_chunk", st)
    return first.value, length.value


def read_matrices(path, count: int, n: int, fmt: str = "complex64", rank: int = 0,
                  nodes: int = 1, device="cuda", stream: torch.cuda.Stream | None = None):
    """Read this rank's run of the reference's raw matrix file onto the
    device.  Returns a complex128 tensor: [m, n, n] when the run holds whole
    matrices starting at a matrix boundary, else the flat run of complex
    values (a run of odd length keeps its last real part in an extra slot)."""
    first, length = file_chunk(count, n, rank, nodes)
    flat = torch.empty(max(1, (length + 1) // 2), dtype=torch.complex128, device=device)
    got = _i64()
    st = _lib.lufqhe_read_matrices(str(path).encode(), FILE_FORMATS[fmt], count, n, rank, nodes,
                                   _ptr(flat), 2 * flat.numel(), ctypes.byref(got),
                                   _stream_handle(stream))
    if st != SUCCESS:
        raise LufqheError("lufqhe_read_matrices", st)
    vals = got.value
    el = n * n * 2
    if first % el == 0 and vals % el == 0:
        return flat[: vals // 2].view(vals // el, n, n)
    return flat[: (vals + 1) // 2]


MIXED_MODES = {"graph": 0, "streams": 1, "plain": 2, "persistent": 3}
SOURCES = {"matrix": 0, "fqhe": 1}


class MixedPlan:
    """A mixed-size batch plan (lufqhe_mixed_plan_*): matrix m has side ns[m]
    and starts at complex element offs[m] of one flat buffer."""

    def __init__(self, ns, offs, policy: str = "crout_ref", nstreams: int = 4):
        import numpy as np
        self.ns = np.ascontiguousarray(ns, dtype=np.int32)
        self.offs = np.ascontiguousarray(offs, dtype=np.int64)
        self.count = len(self.ns)
        if self.offs.shape != self.ns.shape:
            raise ValueError("ns and offs must have the same length")
        # complex elements of the flat buffer the plan reads, per source
        self._need = {"matrix": int((self.offs + self.ns.astype(np.int64) ** 2).max(initial=0)),
                      "fqhe": int((self.offs + self.ns).max(initial=0))}
        self._h = _vp()
        st = load().lufqhe_mixed_plan_create(self.ns.ctypes.data, self.offs.ctypes.data,
                                             self.count, POLICIES[policy], nstreams,
                                             ctypes.byref(self._h))
        if st != SUCCESS:
            raise LufqheError("lufqhe_mixed_plan_create", st)

    def buckets(self):
        """[(n, count, stream)] in launch order."""
        import numpy as np
        nb = _lib.lufqhe_mixed_plan_buckets(self._h, None, None, None, 0)
        n = np.empty(nb, np.int32)
        c = np.empty(nb, np.int64)
        s = np.empty(nb, np.int32)
        _lib.lufqhe_mixed_plan_buckets(self._h, n.ctypes.data, c.ctypes.data, s.ctypes.data, nb)
        return list(zip(n.tolist(), c.tolist(), s.tolist()))

    def execute(self, flat: torch.Tensor, out: DetResult, mode: str = "graph",
                stream: torch.cuda.Stream | None = None, source: str = "matrix") -> DetResult:
        """source "matrix": matrix m at flat[offs[m]:]; "fqhe": its ns[m]
        particle positions there, the FQHE matrix built on the device."""
        if flat.dtype != torch.complex128 or not flat.is_cuda or not flat.is_contiguous():
            raise ValueError("flat must be a contiguous complex128 CUDA tensor")
        if flat.numel() < self._need[source]:
            raise ValueError(f"flat holds {flat.numel()} elements; the plan reads "
                             f"{self._need[source]} for source {source!r}")
        out.check(self.count, 1, flat.device)
        st = _lib.lufqhe_mixed_execute(self._h, _ptr(flat), _ptr(out.mant), _ptr(out.exp),
                                       _ptr(out.logabs), _ptr(out.phase), _ptr(out.info),
                                       SOURCES[source], MIXED_MODES[mode],
                                       _stream_handle(stream))
        if st != SUCCESS:
            raise LufqheError("lufqhe_mixed_execute", st)
        return out

    def close(self) -> None:
        if self._h:
            _lib.lufqhe_mixed_plan_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def logabs_phase_from(mant_re, mant_im, exp):
    """log|det| and arg det from (mant, exp) -- host helper for checks."""
    m = math.hypot(mant_re, mant_im)
    if m == 0.0:
        return -math.inf, 0.0
    return math.log(m) + exp * math.log(2.0), math.atan2(mant_im, mant_re)
