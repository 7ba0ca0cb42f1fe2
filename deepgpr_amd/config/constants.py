"""Physical, numerical and ABI constants of the FDTD hot path.

The values form part of the drop-in contract with the reference package
(``src/DeepGPR/config/constants.py:21-200`` and ``src/DeepGPR/lib/deepgpr.h``):
the storage-type bit layout and the physical constants must be identical, so
histories and arguments mean the same thing on both sides.
"""
from __future__ import annotations

from typing import Dict, Final, Mapping, Tuple

# CODATA 2018, SI units (deepgpr.h:21-24)
SPEED_OF_LIGHT: Final[float] = 299792458.0
VACUUM_PERMEABILITY: Final[float] = 1.25663706212e-06
VACUUM_PERMITTIVITY: Final[float] = 8.8541878128e-12

SUPPORTED_FDTD_ORDERS: Final[Tuple[int, ...]] = (2, 4, 8)

# Sum of |c_r| of the staggered stencil of each order; it scales the CFL limit.
FDTD_STENCIL_COEFFICIENT_SUMS: Final[Mapping[int, float]] = {
    2: 1.0,
    4: 9.0 / 8.0 + 1.0 / 24.0,
    8: 1225.0 / 1024.0 + 245.0 / 3072.0 + 49.0 / 5120.0 + 5.0 / 7168.0,
}

SUPPORTED_GRADIENT_MODES: Final[Tuple[int, ...]] = (2, 3)
FIELD_COMPONENTS: Final[Tuple[int, ...]] = (0, 1, 2)
PEC_CONDUCTIVITY_THRESHOLD: Final[float] = 100.0
MIN_RELATIVE_PERMITTIVITY: Final[float] = 1.0

CPML_SCALING_PROFILES: Final[Dict[str, int]] = {
    "constant": 0, "linear": 1, "quadratic": 2, "cubic": 3, "quartic": 4,
    "quintic": 5, "sextic": 6, "septic": 7, "octic": 8,
}
CPML_SIGMA_MAX_FACTOR: Final[float] = 0.8
CPML_POLE_COUNT: Final[int] = 1
CPML_COEFFICIENTS_PER_CELL: Final[int] = 4
CPML_PHI_TENSORS_PER_FACE: Final[int] = 4
CPML_FACE_COUNT: Final[int] = 6
CPML_PHI_TENSOR_COUNT: Final[int] = CPML_FACE_COUNT * CPML_PHI_TENSORS_PER_FACE
CPML_FACE_NAMES: Final[Tuple[str, ...]] = ("x0", "xm", "y0", "ym", "z0", "zm")

SOURCE_SPECTRUM_CUTOFF_FRACTION: Final[float] = 3.0e-3

NATIVE_ABI_VERSION: Final[int] = 6
CUDA_MAX_SHOTS_PER_CALL: Final[int] = 65535
INT32_MAX: Final[int] = 2**31 - 1

# storage_type: bits 0-3 kind, 4-7 flags, 8-25 INT8 block, 26 no history,
# 27-28 FP16/BF16 conversion backend, 29-30 INT8 reduction backend.
WAVEFIELD_STORAGE_FLOAT32: Final[int] = 0
WAVEFIELD_STORAGE_FLOAT16: Final[int] = 1
WAVEFIELD_STORAGE_BFLOAT16: Final[int] = 2
WAVEFIELD_STORAGE_INT8: Final[int] = 3
WAVEFIELD_STORAGE_KIND_MASK: Final[int] = 0x0F
WAVEFIELD_RHS_FROM_E: Final[int] = 1 << 4
WAVEFIELD_PHYSICAL_HISTORY: Final[int] = 1 << 5
SOLVER_TM2D: Final[int] = 1 << 6
INT8_BLOCK_X_SHIFT: Final[int] = 8
INT8_BLOCK_Y_SHIFT: Final[int] = 14
INT8_BLOCK_Z_SHIFT: Final[int] = 20
INT8_BLOCK_DIM_MASK: Final[int] = 0x3F
WAVEFIELD_HISTORY_DISABLED: Final[int] = 1 << 26
WAVEFIELD_CONVERSION_SHIFT: Final[int] = 27
WAVEFIELD_CONVERSION_BACKENDS: Final[Mapping[str, int]] = {
    "legacy": 0, "native_scalar": 1, "native_vec2": 2,
}
INT8_REDUCTION_SHIFT: Final[int] = 29
INT8_REDUCTION_BACKENDS: Final[Mapping[str, int]] = {
    "current": 0, "cub_block": 1, "warp_shuffle": 2,
}
INT8_MAX_BLOCK_VOLUME: Final[int] = 256
INT8_SPECIALISED_REDUCTION_VOLUME: Final[int] = 64
INT8_DEFAULT_BLOCK_2D: Final[Tuple[int, int, int]] = (8, 8, 1)
INT8_DEFAULT_BLOCK_3D: Final[Tuple[int, int, int]] = (4, 4, 4)
INT8_SCALE_ALIGNMENT_BYTES: Final[int] = 4
FLOAT32_BYTES: Final[int] = 4

# CPML auxiliary arrays (native order) that only couple Ex, Ey and Hz in 2D.
TM2D_INACTIVE_PHI_INDICES: Final[Tuple[int, ...]] = (0, 3, 4, 7, 8, 11, 12, 15)

WAVEFIELD_HISTORY_REGIONS: Final[Tuple[str, ...]] = ("extended", "physical")
WAVEFIELD_RHS_HISTORY_MODES: Final[Tuple[str, ...]] = ("auto", "stored", "reconstructed")

# Gradient accumulators of this build (keyword-only extension of compute()).
GRADIENT_ACCUMULATORS: Final[Tuple[str, ...]] = ("float32", "float64")
