"""Wavefield-history options, the storage_type bit field and the packed INT8
history layout.

What is interface, and therefore the reference's (src/DeepGPR/solver/
storage.py:86-460, config/constants.py:37-200): the option names, values and
messages, the storage_type bit positions read by the native library, the
frame shapes, and the packed INT8 byte layout (int8 payload of
(comp, t, shot, x, y, z) padded to 4 bytes, then FP32 tile scales of
(comp, t, shot, bx, by, bz)). The code is this build's: options are parsed by
one table-driven `_Choice`, the layout is a small value object, the bit field
is assembled from a list of (condition, bits) pairs, and the INT8 decoder
gathers each value's tile scale by index instead of expanding the scale
array.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Dict, Mapping, NamedTuple, Optional, Sequence, Tuple

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


# ----------------------------------------------------------------------------- options


@dataclass(frozen=True)
class _Choice:
    """A case-insensitive string option: a non-string raises TypeError, an
    unknown value ValueError, both with the option's message; `none_as` is
    the value None stands for (None: not accepted)."""

    values: Tuple[str, ...]
    message: str
    none_as: Optional[str] = None

    def __call__(self, value: Any) -> str:
        if value is None and self.none_as is not None:
            return self.none_as
        if not isinstance(value, str):
            raise TypeError(self.message)
        key = value.lower()
        if key not in self.values:
            raise ValueError(self.message)
        return key


normalize_wavefield_compression = _Choice(
    ("none", "int8", "zfp"), "wavefield_compression must be 'none', 'int8', or 'zfp'.", none_as="none")
normalize_wavefield_conversion_backend = _Choice(
    ("auto", *WAVEFIELD_CONVERSION_BACKENDS),
    "wavefield_conversion_backend must be 'auto', 'legacy', 'native_scalar', or 'native_vec2'.")
normalize_int8_reduction_backend = _Choice(
    ("auto", *INT8_REDUCTION_BACKENDS),
    "int8_reduction_backend must be 'auto', 'current', 'cub_block', or 'warp_shuffle'.")
normalize_wavefield_rhs_history = _Choice(
    WAVEFIELD_RHS_HISTORY_MODES, "wavefield_rhs_history must be 'auto', 'stored', or 'reconstructed'.")
normalize_wavefield_history_region = _Choice(
    WAVEFIELD_HISTORY_REGIONS, "wavefield_history_region must be 'extended' or 'physical'.")


def normalize_wavefield_storage_dtype(value: Any) -> torch.dtype:
    """float32 / float16 / bfloat16, as a torch dtype or one of its names."""
    dtype = STORAGE_DTYPE_ALIASES.get(value.lower()) if isinstance(value, str) else value
    if dtype not in WAVEFIELD_STORAGE_TYPES:
        raise ValueError("wavefield_storage_dtype must be float32, float16, or bfloat16.")
    return dtype


# ----------------------------------------------------------------------------- saved region


def _tile_span(cells: int, lo: int, hi: int, tile: int) -> Tuple[int, int]:
    """(start, length) of [lo, cells - hi) widened outward to whole tiles of
    the full-grid tiling (tile boundaries at multiples of `tile`), clipped to
    the grid."""
    start = lo - lo % tile
    stop = min(cells, ((cells - hi + tile - 1) // tile) * tile)
    return start, stop - start


def history_box(nx: int, ny: int, nz: int, pml: Sequence[int], region: str,
                block_size: Optional[Sequence[int]] = None):
    """``((x0, y0, z0), (hx, hy, hz))`` of one saved frame: the whole
    extended grid, or the physical region (the grid without the CPML faces)
    widened to whole INT8 tiles when a tile size is given."""
    if region != "physical":
        return (0, 0, 0), (nx, ny, nz)
    tiles = (1, 1, 1) if block_size is None else tuple(int(t) for t in block_size)
    spans = [_tile_span(n, int(pml[2 * a]), int(pml[2 * a + 1]), tiles[a]) for a, n in enumerate((nx, ny, nz))]
    return tuple(s for s, _ in spans), tuple(n for _, n in spans)


def history_spatial_shape(nx, ny, nz, pml, region, block_size=None):
    return history_box(nx, ny, nz, pml, region, block_size)[1]


def resolve_rhs_reconstruction(requested: str, *, sampling_interval: int, stores_model_history: bool,
                               library_supported: bool, lossless_storage: bool = True) -> bool:
    """Whether the history keeps E^n only and the adjoint rebuilds
    R^n = (E^(n+1) - ca E^n) / cb from consecutive frames (possible only
    with every step saved). 'auto' does so for lossless storage;
    'reconstructed' demands it (an error where impossible)."""
    if requested == "stored":
        return False
    every_step = sampling_interval == 1
    if requested == "auto":
        return every_step and lossless_storage and library_supported and stores_model_history
    if not every_step:
        raise ValueError("wavefield_rhs_history='reconstructed' requires model_gradient_sampling_interval=1 "
                         "(R^n is rebuilt from E^n and E^(n+1)).")
    if stores_model_history and not library_supported:
        raise NativeLibraryError("The loaded native library cannot rebuild R^n from the E history.")
    return stores_model_history


# ----------------------------------------------------------------------------- INT8 blocks and bits


def default_compression_block_size(spatial_mode: int) -> BlockSize:
    return INT8_DEFAULT_BLOCK_2D if spatial_mode == 2 else INT8_DEFAULT_BLOCK_3D


def normalize_compression_block_size(value: Any, spatial_mode: int) -> BlockSize:
    """INT8 tile (bx, by, bz): two values for a 2D model (bz = 1), three for
    3D; each in [1, 63], volume a power of two of at most 256."""
    if value is None:
        return default_compression_block_size(spatial_mode)
    if isinstance(value, int):
        raise TypeError("wavefield_compression_block_size must be a 2D or 3D sequence, not a scalar.")
    try:
        dims = [int(v) for v in value]
    except (TypeError, ValueError) as exc:
        raise TypeError("wavefield_compression_block_size must be a sequence of integers.") from exc
    want = 2 if spatial_mode == 2 else 3
    if len(dims) != want:
        raise ValueError(f"wavefield_compression_block_size must contain {want} values for a {spatial_mode}D model.")
    block = tuple(dims + [1] * (3 - want))
    if min(block) < 1 or max(block) > INT8_BLOCK_DIM_MASK:
        raise ValueError("Each INT8 compression block dimension must be in [1, 63].")
    volume = math.prod(block)
    if volume > INT8_MAX_BLOCK_VOLUME or bin(volume).count("1") != 1:
        raise ValueError("The INT8 compression block volume must be a power of two no larger than 256.")
    return block  # type: ignore[return-value]


def encode_int8_storage_type(block_size: BlockSize, reduction_backend: str = "current") -> int:
    fields = zip(block_size, (INT8_BLOCK_X_SHIFT, INT8_BLOCK_Y_SHIFT, INT8_BLOCK_Z_SHIFT))
    bits = WAVEFIELD_STORAGE_INT8 | INT8_REDUCTION_BACKENDS[reduction_backend] << INT8_REDUCTION_SHIFT
    for dim, shift in fields:
        bits |= int(dim) << shift
    return bits


def encode_storage_type(config: WavefieldStorageConfig, save_wavefield_history: bool, rhs_from_e: bool = False,
                        physical_history: bool = False, tm2d: bool = False) -> int:
    """The storage_type argument of the native forward/backward."""
    if config.compression == "int8":
        kind = encode_int8_storage_type(config.block_size, config.int8_reduction_backend)  # type: ignore[arg-type]
    else:
        kind = WAVEFIELD_STORAGE_TYPES[config.dtype] | (
            WAVEFIELD_CONVERSION_BACKENDS[config.conversion_backend] << WAVEFIELD_CONVERSION_SHIFT)
    flags = (
        (not save_wavefield_history, WAVEFIELD_HISTORY_DISABLED),
        (save_wavefield_history and rhs_from_e, WAVEFIELD_RHS_FROM_E),
        (save_wavefield_history and physical_history, WAVEFIELD_PHYSICAL_HISTORY),
        (tm2d, SOLVER_TM2D),
    )
    for on, bits in flags:
        if on:
            kind |= bits
    return kind


def saved_history_shape(mode: int, nt_saved: int, nstep: int, nx: int, ny: int, nz: int) -> Tuple[int, ...]:
    frame = (nt_saved, nstep, nx, ny, nz)
    return (3, *frame) if mode == 3 else frame


def saved_time_steps(nt: int, sampling_interval: int) -> int:
    return -(-nt // sampling_interval)


# ----------------------------------------------------------------------------- packed INT8 layout


@dataclass(frozen=True)
class _Int8Layout:
    components: int
    nt_saved: int
    shots: int
    spatial: Tuple[int, int, int]
    block: BlockSize

    @property
    def tiles(self) -> Tuple[int, int, int]:
        return tuple(-(-n // b) for n, b in zip(self.spatial, self.block))  # type: ignore[return-value]

    @property
    def frames(self) -> int:
        return self.components * self.nt_saved * self.shots

    @property
    def payload(self) -> int:
        return self.frames * math.prod(self.spatial)

    @property
    def scales_offset(self) -> int:
        a = INT8_SCALE_ALIGNMENT_BYTES
        return -(-self.payload // a) * a

    @property
    def scale_count(self) -> int:
        return self.frames * math.prod(self.tiles)

    def as_dict(self) -> Mapping[str, Any]:
        return {"components": self.components, "nt_saved": self.nt_saved, "shots": self.shots,
                "spatial_shape": self.spatial, "blocks_xyz": self.tiles, "blocks_per_shot": math.prod(self.tiles),
                "q_count": self.payload, "q_bytes_aligned": self.scales_offset, "scale_count": self.scale_count,
                "packed_bytes": self.scales_offset + self.scale_count * FLOAT32_BYTES}


def _layout(shape: Sequence[int], block_size: Sequence[int]) -> _Int8Layout:
    if len(shape) not in (5, 6):
        raise ValueError("Wavefield history shape must have five or six dimensions.")
    comps = 1 if len(shape) == 5 else int(shape[0])
    nt_saved, shots, nx, ny, nz = (int(v) for v in shape[-5:])
    return _Int8Layout(comps, nt_saved, shots, (nx, ny, nz), tuple(int(b) for b in block_size))  # type: ignore


def int8_history_layout(shape: Sequence[int], block_size: BlockSize) -> Dict[str, Any]:
    """Sizes and offsets of a packed INT8 history of `shape` (five or six
    dimensions, as saved_history_shape) with tiles `block_size`."""
    return dict(_layout(shape, block_size).as_dict())


def decompress_wavefield_history(packed: torch.Tensor, shape: Sequence[int], block_size: Sequence[int]) -> torch.Tensor:
    """float32 reconstruction of a packed INT8 history (diagnostics only):
    value = q * scale of its tile."""
    block = tuple(block_size) + ((1,) if len(block_size) == 2 else ())
    if len(block) != 3:
        raise ValueError("block_size must contain two or three values.")
    lay = _layout(shape, block)
    info = lay.as_dict()
    if packed.dtype != torch.int8 or packed.ndim != 1 or not packed.is_contiguous():
        raise TypeError("packed must be a contiguous one-dimensional torch.int8 tensor.")
    if packed.numel() != info["packed_bytes"]:
        raise ValueError(f"packed contains {packed.numel()} bytes; expected {info['packed_bytes']}.")
    nx, ny, nz = lay.spatial
    tx, ty, tz = lay.tiles
    q = packed[:lay.payload].reshape(lay.frames, nx, ny, nz)
    scales = packed[lay.scales_offset:].view(torch.float32).reshape(lay.frames, tx * ty * tz)
    dev = packed.device
    tile_of = ((torch.arange(nx, device=dev) // block[0])[:, None, None] * ty
               + (torch.arange(ny, device=dev) // block[1])[None, :, None]) * tz \
        + (torch.arange(nz, device=dev) // block[2])[None, None, :]
    out = q.to(torch.float32) * scales[:, tile_of.reshape(-1)].reshape(lay.frames, nx, ny, nz)
    return out.reshape(tuple(int(v) for v in shape))
