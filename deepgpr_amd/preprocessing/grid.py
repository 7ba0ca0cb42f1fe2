"""Grid spacing and PML thickness normalisation (src/DeepGPR/preprocessing/grid.py:17-116)."""
from __future__ import annotations

import math
import numbers
from typing import Any, Sequence, Tuple, Union

import torch

from ..config.constants import INT32_MAX

GridSpacing = Tuple[float, float, float]
GridSpacingLike = Union[float, Sequence[float], torch.Tensor]


def normalize_grid_spacing(value: GridSpacingLike) -> GridSpacing:
    """``(dx, dy, dz)`` from a scalar, a 3-sequence or a 1/3-element tensor."""
    if isinstance(value, torch.Tensor):
        flat = value.detach().cpu().reshape(-1)
        if flat.numel() == 1:
            items = [flat.item()] * 3
        elif flat.numel() == 3:
            items = flat.tolist()
        else:
            raise ValueError("dx tensor must contain either one or three values.")
    elif isinstance(value, numbers.Real) and not isinstance(value, bool):
        items = [value] * 3
    elif isinstance(value, (list, tuple)):
        if len(value) != 3:
            raise ValueError("dx list or tuple must contain exactly three values.")
        items = list(value)
    else:
        raise TypeError("dx must be a positive scalar or a three-value list, tuple, or tensor.")
    try:
        spacing = tuple(float(v) for v in items)
    except (TypeError, ValueError) as exc:
        raise TypeError("dx, dy, and dz must be finite positive scalars.") from exc
    if not all(math.isfinite(v) and v > 0.0 for v in spacing):
        raise ValueError("dx, dy, and dz must be finite positive scalars.")
    return spacing  # type: ignore[return-value]


def pmlthick_revert(p: Any, er: torch.Tensor) -> torch.Tensor:
    """Six face thicknesses ``[x0, xm, y0, ym, z0, zm]`` as an int32 CPU tensor.

    An int applies to every active face (z faces are 0 in 2D); four values are
    ``[x0, xm, y0, ym]``; six values are used as given.
    """
    if isinstance(p, bool):
        raise TypeError("PML thickness must contain integer values, not bool.")
    if isinstance(p, numbers.Integral):
        n = int(p)
        values = [n, n, n, n, 0, 0] if er.shape[2] == 1 else [n] * 6
    elif isinstance(p, (list, tuple)) or isinstance(p, torch.Tensor):
        if isinstance(p, torch.Tensor):
            if p.ndim != 1:
                raise ValueError("PML thickness tensor must be one-dimensional.")
            seq = p.detach().cpu().tolist()
        else:
            seq = list(p)
        if len(seq) == 4:
            values = [*seq, 0, 0]
        elif len(seq) == 6:
            values = seq
        else:
            raise ValueError(f"Unsupported PML length: {len(seq)}. Must be 4 or 6.")
    else:
        raise TypeError(f"Unsupported PML thickness type: {type(p)}")
    out = []
    for v in values:
        if isinstance(v, bool):
            raise TypeError("PML thickness must contain integer values, not bool.")
        try:
            f = float(v)
        except (TypeError, ValueError, OverflowError) as exc:
            raise TypeError("PML thickness must contain finite integer values.") from exc
        if not math.isfinite(f) or not f.is_integer():
            raise ValueError("PML thickness must contain finite integer values.")
        if f < 0 or f > INT32_MAX:
            raise ValueError("PML thickness values must fit in non-negative int32.")
        out.append(int(f))
    return torch.tensor(out, dtype=torch.int32)
