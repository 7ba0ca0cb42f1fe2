"""Default values of the public API (src/DeepGPR/config/defaults.py:13-44)."""
from __future__ import annotations

from typing import Final

DEFAULT_PML_THICKNESS: Final[int] = 10
DEFAULT_FDTD_ORDER: Final[int] = 2
DEFAULT_GRADIENT_MODE: Final[int] = 2
DEFAULT_SOURCE_COMPONENT: Final[int] = 2
DEFAULT_RECEIVER_COMPONENT: Final[int] = 2
DEFAULT_GRADIENT_SAMPLING_INTERVAL: Final[int] = 1
DEFAULT_WAVEFIELD_COMPRESSION: Final[str] = "none"
DEFAULT_WAVEFIELD_RHS_HISTORY: Final[str] = "auto"
DEFAULT_WAVEFIELD_HISTORY_REGION: Final[str] = "extended"
DEFAULT_CONVERSION_BACKEND: Final[str] = "auto"
DEFAULT_INT8_REDUCTION_BACKEND: Final[str] = "auto"
# Time-summed material gradients: "float64" accumulators on the device (this
# build's default) or "float32" (bitwise the reference's accumulation).
DEFAULT_GRADIENT_ACCUMULATOR: Final[str] = "float64"
LOGGER_NAME: Final[str] = "DeepGPR"
