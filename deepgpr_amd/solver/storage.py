"""Wavefield-history options, the storage_type bit field and the packed INT8
layout (src/DeepGPR/solver/storage.py:86-460)."""
from __future__ import annotations

import math
from typing import Any, Dict, NamedTuple, Optional, Sequence, Tuple

import torch

from ..config.constants import (
    FLOAT32_BYTES, INT8_BLOCK_DIM_MASK, INT8_BLOCK_X_SHIFT, INT8_BLOCK_Y_SHIFT, INT8_BLOCK_Z_SHIFT,
    INT8_DEFAULT_BLOCK_2D, INT8_DEFAULT_BLOCK_3D, INT8_MAX_BLOCK_VOLUME, INT8_REDUCTION_BACKENDS,
    INT8_REDUCTION_SHIFT, INT8_SCALE_ALIGNMENT_BYTES, SOLVER_TM2D, WAVEFIELD_CONVERSION_BACKENDS,
    WAVEFIELD_CONVERSION_SHIFT, WAVEFIELD_HISTORY_DISABLED, WAVEFIELD_HISTORY_REGIONS,
    WAVEFIELD_PHYSICAL_HISTORY, WAVEFIELD_RHS_FROM_E, WAVEFIELD_RHS_HISTORY_MODES,
    WAVEFIELD_STORAGE_BFLOAT16, WAVEFIELD_STORAGE_FLOAT16, WAVEFIELD_STORAGE_FLOAT32, WAVEFIELD_STORAGE_INT8,
)
from ..utils.exceptions import NativeLibraryError

BlockSize = Tuple[int, int, int]

WAVEFIELD_STORAGE_TYPES: Dict[torch.dtype, int] = {
    torch.float32: WAVEFIELD_STORAGE_FLOAT32,
    torch.float16: WAVEFIELD_STORAGE_FLOAT16,
    torch.bfloat16: WAVEFIELD_STORAGE_BFLOAT16,
}
STORAGE_DTYPE_ALIASES: Dict[str, torch.dtype] = {
    "float32": torch.float32, "fp32": torch.float32, "float": torch.float32,
    "float16": torch.float16, "fp16": torch.float16, "half": torch.float16,
    "bfloat16": torch.bfloat16, "bf16": torch.bfloat16,
}


class WavefieldStorageConfig(NamedTuple):
    dtype: torch.dtype
    compression: str
    block_size: Optional[BlockSize]
    conversion_backend: str
    int8_reduction_backend: str


def normalize_wavefield_storage_dtype(value: Any) -> torch.dtype:
    if isinstance(value, str):
        value = STORAGE_DTYPE_ALIASES.get(value.lower())
    if value not in WAVEFIELD_STORAGE_TYPES:
        raise ValueError("wavefield_storage_dtype must be float32, float16, or bfloat16.")
    return value


def normalize_wavefield_compression(value: Any) -> str:
    if value is None:
        return "none"
    if not isinstance(value, str):
        raise TypeError("wavefield_compression must be 'none', 'int8', or 'zfp'.")
    value = value.lower()
    if value not in ("none", "int8", "zfp"):
        raise ValueError("wavefield_compression must be 'none', 'int8', or 'zfp'.")
    return value


def _choice(value: Any, allowed, message: str) -> str:
    if not isinstance(value, str):
        raise TypeError(message)
    value = value.lower()
    if value not in allowed:
        raise ValueError(message)
    return value


def normalize_wavefield_conversion_backend(value: Any) -> str:
    return _choice(value, ("auto", *WAVEFIELD_CONVERSION_BACKENDS),
                   "wavefield_conversion_backend must be 'auto', 'legacy', 'native_scalar', or 'native_vec2'.")


def normalize_int8_reduction_backend(value: Any) -> str:
    return _choice(value, ("auto", *INT8_REDUCTION_BACKENDS),
                   "int8_reduction_backend must be 'auto', 'current', 'cub_block', or 'warp_shuffle'.")


def normalize_wavefield_rhs_history(value: Any) -> str:
    return _choice(value, WAVEFIELD_RHS_HISTORY_MODES,
                   "wavefield_rhs_history must be 'auto', 'stored', or 'reconstructed'.")


def normalize_wavefield_history_region(value: Any) -> str:
    return _choice(value, WAVEFIELD_HISTORY_REGIONS, "wavefield_history_region must be 'extended' or 'physical'.")


def history_box(nx: int, ny: int, nz: int, pml: Sequence[int], region: str,
                block_size: Optional[Sequence[int]] = None):
    """``((x0, y0, z0), (hx, hy, hz))`` of one saved frame; the physical box is
    widened to whole INT8 tiles of the full-grid tiling."""
    if region != "physical":
        return (0, 0, 0), (nx, ny, nz)
    tiles = tuple(block_size) if block_size is not None else (1, 1, 1)
    origin, extent = [], []
    for cells, lo, hi, t in zip((nx, ny, nz), pml[0::2], pml[1::2], tiles):
        begin = (int(lo) // t) * t
        end = min(cells, -(-(cells - int(hi)) // t) * t)
        origin.append(begin)
        extent.append(end - begin)
    return tuple(origin), tuple(extent)


def history_spatial_shape(nx, ny, nz, pml, region, block_size=None):
    return history_box(nx, ny, nz, pml, region, block_size)[1]


def resolve_rhs_reconstruction(requested: str, *, sampling_interval: int, stores_model_history: bool,
                               library_supported: bool, lossless_storage: bool = True) -> bool:
    """Whether R^n is rebuilt from consecutive E frames (E-only history)."""
    if requested == "stored":
        return False
    if requested == "reconstructed":
        if sampling_interval != 1:
            raise ValueError("wavefield_rhs_history='reconstructed' requires model_gradient_sampling_interval=1 "
                             "(R^n is rebuilt from E^n and E^(n+1)).")
        if not stores_model_history:
            return False
        if not library_supported:
            raise NativeLibraryError("The loaded native library cannot rebuild R^n from the E history.")
        return True
    return sampling_interval == 1 and lossless_storage and library_supported and stores_model_history


def default_compression_block_size(spatial_mode: int) -> BlockSize:
    return INT8_DEFAULT_BLOCK_2D if spatial_mode == 2 else INT8_DEFAULT_BLOCK_3D


def normalize_compression_block_size(value: Any, spatial_mode: int) -> BlockSize:
    if value is None:
        return default_compression_block_size(spatial_mode)
    if isinstance(value, int):
        raise TypeError("wavefield_compression_block_size must be a 2D or 3D sequence, not a scalar.")
    try:
        block = tuple(int(v) for v in value)
    except (TypeError, ValueError) as exc:
        raise TypeError("wavefield_compression_block_size must be a sequence of integers.") from exc
    dims = 2 if spatial_mode == 2 else 3
    if len(block) != dims:
        raise ValueError(f"wavefield_compression_block_size must contain {dims} values for a {spatial_mode}D model.")
    if spatial_mode == 2:
        block = (*block, 1)
    if any(b < 1 or b > INT8_BLOCK_DIM_MASK for b in block):
        raise ValueError("Each INT8 compression block dimension must be in [1, 63].")
    vol = math.prod(block)
    if vol > INT8_MAX_BLOCK_VOLUME or vol & (vol - 1):
        raise ValueError("The INT8 compression block volume must be a power of two no larger than 256.")
    return block  # type: ignore[return-value]


def encode_int8_storage_type(block_size: BlockSize, reduction_backend: str = "current") -> int:
    bx, by, bz = block_size
    return (WAVEFIELD_STORAGE_INT8 | (bx << INT8_BLOCK_X_SHIFT) | (by << INT8_BLOCK_Y_SHIFT)
            | (bz << INT8_BLOCK_Z_SHIFT) | (INT8_REDUCTION_BACKENDS[reduction_backend] << INT8_REDUCTION_SHIFT))


def encode_storage_type(config: WavefieldStorageConfig, save_wavefield_history: bool, rhs_from_e: bool = False,
                        physical_history: bool = False, tm2d: bool = False) -> int:
    if config.compression == "int8":
        st = encode_int8_storage_type(config.block_size, config.int8_reduction_backend)  # type: ignore[arg-type]
    else:
        st = WAVEFIELD_STORAGE_TYPES[config.dtype] | (
            WAVEFIELD_CONVERSION_BACKENDS[config.conversion_backend] << WAVEFIELD_CONVERSION_SHIFT)
    if not save_wavefield_history:
        st |= WAVEFIELD_HISTORY_DISABLED
    else:
        if rhs_from_e:
            st |= WAVEFIELD_RHS_FROM_E
        if physical_history:
            st |= WAVEFIELD_PHYSICAL_HISTORY
    if tm2d:
        st |= SOLVER_TM2D
    return st


def saved_history_shape(mode: int, nt_saved: int, nstep: int, nx: int, ny: int, nz: int) -> Tuple[int, ...]:
    return (3, nt_saved, nstep, nx, ny, nz) if mode == 3 else (nt_saved, nstep, nx, ny, nz)


def saved_time_steps(nt: int, sampling_interval: int) -> int:
    return (nt + sampling_interval - 1) // sampling_interval


def int8_history_layout(shape: Sequence[int], block_size: BlockSize) -> Dict[str, Any]:
    """Packed INT8 layout: int8 payload (comp, t, shot, x, y, z) padded to 4
    bytes, then FP32 scales (comp, t, shot, bx, by, bz)."""
    if len(shape) == 5:
        comps = 1
        nt_saved, shots, nx, ny, nz = shape
    elif len(shape) == 6:
        comps, nt_saved, shots, nx, ny, nz = shape
    else:
        raise ValueError("Wavefield history shape must have five or six dimensions.")
    bx, by, bz = block_size
    blocks = ((nx + bx - 1) // bx, (ny + by - 1) // by, (nz + bz - 1) // bz)
    per_shot = math.prod(blocks)
    q = comps * nt_saved * shots * nx * ny * nz
    mask = INT8_SCALE_ALIGNMENT_BYTES - 1
    q_aligned = (q + mask) & ~mask
    scales = comps * nt_saved * shots * per_shot
    return {"components": comps, "nt_saved": nt_saved, "shots": shots, "spatial_shape": (nx, ny, nz),
            "blocks_xyz": blocks, "blocks_per_shot": per_shot, "q_count": q, "q_bytes_aligned": q_aligned,
            "scale_count": scales, "packed_bytes": q_aligned + scales * FLOAT32_BYTES}


def decompress_wavefield_history(packed: torch.Tensor, shape: Sequence[int], block_size: Sequence[int]) -> torch.Tensor:
    """float32 reconstruction of a packed INT8 history (diagnostics only)."""
    block = tuple(block_size)
    if len(block) == 2:
        block = (*block, 1)
    if len(block) != 3:
        raise ValueError("block_size must contain two or three values.")
    lay = int8_history_layout(tuple(shape), block)  # type: ignore[arg-type]
    if packed.dtype != torch.int8 or packed.ndim != 1 or not packed.is_contiguous():
        raise TypeError("packed must be a contiguous one-dimensional torch.int8 tensor.")
    if packed.numel() != lay["packed_bytes"]:
        raise ValueError(f"packed contains {packed.numel()} bytes; expected {lay['packed_bytes']}.")
    c, t, s = lay["components"], lay["nt_saved"], lay["shots"]
    nx, ny, nz = lay["spatial_shape"]
    nbx, nby, nbz = lay["blocks_xyz"]
    q = packed[: lay["q_count"]].reshape(c, t, s, nx, ny, nz)
    sc = packed[lay["q_bytes_aligned"]: lay["packed_bytes"]].view(torch.float32).reshape(c, t, s, nbx, nby, nbz)
    sc = sc.repeat_interleave(block[0], dim=-3).repeat_interleave(block[1], dim=-2)
    sc = sc.repeat_interleave(block[2], dim=-1)[..., :nx, :ny, :nz]
    out = q.to(torch.float32) * sc
    return out.squeeze(0) if len(shape) == 5 else out
