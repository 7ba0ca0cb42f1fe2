"""Errors and logging."""
from .exceptions import (  # noqa: F401
    CFLConditionError, DeepGPRError, NativeLibraryError, NativeLibraryNotFoundError,
    NonFiniteTensorError,
)
