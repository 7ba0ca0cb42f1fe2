"""Marshal tensors into ``struct dgx_solve`` and call the HIP library.

Replaces the reference bridge (src/DeepGPR/native/bridge.py:65-255): the FDTD
order travels in the descriptor instead of a process-global setter, so calls
with different orders never need to be serialised, and the kernels run on
torch's current stream instead of the legacy default stream.
"""
from __future__ import annotations

import ctypes
from typing import Any, Dict, Optional, Sequence

import torch

from ..utils.exceptions import NativeLibraryError
from .abi import DGX_ABI_VERSION, DgxSolve
from .loader import get_library


def _ptr(value: Any) -> Optional[int]:
    if value is None:
        return None
    if isinstance(value, torch.Tensor):
        if value.numel() == 0:
            return None
        if not value.is_contiguous():
            raise ValueError("native arrays must be contiguous")
        return value.data_ptr()
    return int(value)


def _check_device(tensors: Sequence[Any], device: torch.device) -> None:
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.numel() and t.device != device:
            raise ValueError(f"tensor on {t.device} passed to a solve on {device}")


def build_solve(args: Dict[str, Any]) -> DgxSolve:
    """Fill a DgxSolve from named values (tensors or ints/floats)."""
    s = DgxSolve()
    s.abi_version = DGX_ABI_VERSION
    for name, ctype in DgxSolve._fields_:
        if name == "abi_version" or name not in args:
            continue
        value = args[name]
        if name in ("E", "H", "update_coeff", "phi", "pml_coeff"):
            arr = getattr(s, name)
            for k, v in enumerate(value):
                arr[k] = _ptr(v)
        elif name == "pml":
            for k, v in enumerate(value):
                s.pml[k] = int(v)
        elif ctype is ctypes.c_void_p:
            setattr(s, name, _ptr(value))
        elif ctype is ctypes.c_float:
            setattr(s, name, float(value))
        else:
            setattr(s, name, int(value))
    return s


def invoke(function: str, args: Dict[str, Any], device: torch.device) -> None:
    """Run ``dgx_forward`` or ``dgx_backward`` on ``device``'s current stream."""
    lib = get_library()
    keep = [v for v in args.values() if isinstance(v, torch.Tensor)]
    for v in args.values():
        if isinstance(v, (list, tuple)):
            keep.extend(t for t in v if isinstance(t, torch.Tensor))
    _check_device(keep, device)
    solve = build_solve(args)
    with torch.cuda.device(device):
        stream = torch.cuda.current_stream(device).cuda_stream
        fn = getattr(lib, "dgx_" + function)
        status = fn(ctypes.byref(solve), ctypes.c_void_p(stream))
    del keep
    if status != 0:
        message = lib.dgx_last_error()
        if isinstance(message, bytes):
            message = message.decode("utf-8", "replace")
        raise NativeLibraryError(f"Native DeepGPR {function} failed: {message}")


def scratch_bytes(args: Dict[str, Any], backward: bool) -> int:
    lib = get_library()
    solve = build_solve(args)
    value = int(lib.dgx_scratch_bytes(ctypes.byref(solve), 1 if backward else 0))
    if value < 0:
        message = lib.dgx_last_error()
        raise NativeLibraryError(message.decode() if isinstance(message, bytes) else str(message))
    return value
