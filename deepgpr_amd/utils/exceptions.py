"""Error types. Names and built-in bases match the reference
(src/DeepGPR/utils/exceptions.py:11-30) so existing handlers keep working."""
from __future__ import annotations


class DeepGPRError(Exception):
    """Base class of every error raised by this package."""


class NativeLibraryError(DeepGPRError, RuntimeError):
    """The HIP library is missing a capability, has the wrong ABI, or a call failed."""


class NativeLibraryNotFoundError(DeepGPRError, FileNotFoundError):
    """The HIP library has not been built or cannot be loaded."""


class CFLConditionError(DeepGPRError, ValueError):
    """The time step violates the CFL limit of the staggered scheme."""


class NonFiniteTensorError(DeepGPRError, FloatingPointError):
    """A tensor produced or consumed by the solver holds NaN or Inf."""


__all__ = [
    "CFLConditionError", "DeepGPRError", "NativeLibraryError",
    "NativeLibraryNotFoundError", "NonFiniteTensorError",
]
