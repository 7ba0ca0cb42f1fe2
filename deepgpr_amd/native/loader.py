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
