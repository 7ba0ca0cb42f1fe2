"""Locate and load the in-tree HIP library.

The library is built in place by ``__graft_entry__.build()`` (or
``python -m deepgpr_amd.build``) as ``deepgpr_amd/libdeepgpr_amd.so``. There is
no fallback: a missing or stale library raises, so no call can silently run
anything but the HIP kernels (the reference loader's behaviour for a missing
backend, src/DeepGPR/native/loader.py:174-235).
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

from ..utils.exceptions import NativeLibraryError, NativeLibraryNotFoundError
from .abi import CTYPES_BY_KIND, DGX_ABI_VERSION, NATIVE_FUNCTIONS, REQUIRED_SYMBOLS

_LOCK = threading.Lock()
_LIB: Optional[ctypes.CDLL] = None

LIBRARY_NAME = "libdeepgpr_amd.so"


def library_path() -> str:
    override = os.environ.get("DEEPGPR_AMD_LIBRARY")
    if override:
        return override
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), LIBRARY_NAME)


def _configure(lib: ctypes.CDLL) -> None:
    for name in REQUIRED_SYMBOLS:
        if not hasattr(lib, name):
            raise NativeLibraryError(f"{library_path()} does not export {name}; rebuild it.")
    for name, (params, ret) in NATIVE_FUNCTIONS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = [CTYPES_BY_KIND[kind] for _, kind in params]
        fn.restype = None if ret is None else CTYPES_BY_KIND[ret]
    if lib.dgx_abi_version() != DGX_ABI_VERSION:
        raise NativeLibraryError(
            f"{library_path()} implements dgx ABI {lib.dgx_abi_version()}, expected {DGX_ABI_VERSION}; rebuild it.")
    if lib.deepgpr_abi_version() != 6:
        raise NativeLibraryError("library does not implement the reference ABI 6 export set")


def get_library() -> ctypes.CDLL:
    """Load (once) and return the HIP library."""
    global _LIB
    with _LOCK:
        if _LIB is None:
            path = library_path()
            if not os.path.isfile(path):
                raise NativeLibraryNotFoundError(
                    f"HIP library not found at {path}; build it with __graft_entry__.build() "
                    "or `python -m deepgpr_amd.build`.")
            import torch  # noqa: F401  (load torch's HIP runtime first: one runtime per process)
            try:
                lib = ctypes.CDLL(path)
            except OSError as exc:
                raise NativeLibraryNotFoundError(f"cannot load {path}: {exc}") from exc
            _configure(lib)
            _LIB = lib
        return _LIB


def require_library(device) -> ctypes.CDLL:
    """Return the library for a tensor device; only HIP devices are served."""
    import torch

    device = torch.device(device)
    if device.type != "cuda":
        raise NativeLibraryError(
            "deepgpr_amd runs the FDTD hot path on MI355X (HIP) devices only; "
            f"device {device} has no backend in this build.")
    return get_library()


def library_supports(lib, capability: str) -> bool:
    probe = getattr(lib, capability, None)
    return probe is not None and int(probe()) == 1


# The reference loader's public helpers (src/DeepGPR/native/loader.py:222-293,
# exported at src/DeepGPR/__init__.py:36-41), with the same names, arguments
# and errors: an unknown device kind is a ValueError; the CPU kind, which the
# reference serves with its C backend, has no library here and raises
# NativeLibraryError, as compute() does for a CPU device.

def get_deepgpr_lib(device) -> ctypes.CDLL:
    """The library serving `device` ("cuda", "cuda:N", a torch.device)."""
    import torch

    kind = torch.device(device).type
    if kind not in ("cpu", "cuda"):
        raise ValueError(f"Unsupported DeepGPR device: {device}")
    return require_library(device)


def get_deepgpr_library_path(device) -> str:
    """Filesystem path of the library serving `device`."""
    get_deepgpr_lib(device)
    return library_path()


def clear_cuda_resource_cache(device=None) -> None:
    """Release the library's idle per-device resources (None: the current
    device; an int: that device index). Histories stay in HBM here, so the
    library holds no offload buffers and the call releases nothing; it keeps
    the reference's device checks and error reporting."""
    import torch

    if isinstance(device, int):
        target = torch.device("cuda", device)
    else:
        target = torch.device("cuda" if device is None else device)
    if target.type != "cuda":
        raise ValueError("clear_cuda_resource_cache requires a CUDA device")
    lib = get_deepgpr_lib(target)
    with torch.cuda.device(target):
        lib.deepgpr_clear_last_error()
        lib.deepgpr_clear_cuda_resource_cache()
        message = lib.deepgpr_last_error()
    if message:
        if isinstance(message, bytes):
            message = message.decode("utf-8", "replace")
        raise NativeLibraryError(f"deepgpr_clear_cuda_resource_cache failed: {message}")


def set_library_fdtd_order(lib, fdtd_order: int) -> None:
    """Select the FDTD order of the ABI-6 `forward` / `backward` entry points
    (process-global state in the library; the dgx_* entry points take the
    order in their descriptor and ignore it)."""
    from ..config.constants import SUPPORTED_FDTD_ORDERS

    if fdtd_order not in SUPPORTED_FDTD_ORDERS:
        raise ValueError("fdtd_order must be one of 2, 4, or 8.")
    lib.set_fdtd_order(int(fdtd_order))
