"""Grid spacing and PML thickness normalisation.

The accepted input forms, the exception types and the messages are those of
src/DeepGPR/preprocessing/grid.py:17-116 (callers and the reference's tests
match on them); the implementation is this build's: every accepted form is
first flattened to a list by `_flatten`, expanded to its full length by a
per-length rule table, then converted and checked element by element.
"""
from __future__ import annotations

import math
import numbers
from typing import Any, Callable, Dict, List, Sequence, Tuple, Union

import torch

from ..config.constants import INT32_MAX

GridSpacing = Tuple[float, float, float]
GridSpacingLike = Union[float, Sequence[float], torch.Tensor]

_SPACING_TYPE = "dx must be a positive scalar or a three-value list, tuple, or tensor."
_SPACING_VALUE = "dx, dy, and dz must be finite positive scalars."
_PML_BOOL = "PML thickness must contain integer values, not bool."
_PML_INT = "PML thickness must contain finite integer values."
_PML_RANGE = "PML thickness values must fit in non-negative int32."


def _is_real_scalar(v: Any) -> bool:
    return isinstance(v, numbers.Real) and not isinstance(v, bool)


def _flatten(value: Any) -> Tuple[str, List[Any]]:
    """('scalar' | 'sequence' | 'tensor', items) for the three accepted shapes
    of a multi-valued option; anything else is ('other', [])."""
    if isinstance(value, torch.Tensor):
        return "tensor", value.detach().cpu().reshape(-1).tolist()
    if isinstance(value, (list, tuple)):
        return "sequence", list(value)
    if isinstance(value, numbers.Number) and not isinstance(value, bool):
        return "scalar", [value]
    return "other", []


def normalize_grid_spacing(value: GridSpacingLike) -> GridSpacing:
    """``(dx, dy, dz)`` in metres from a positive scalar (isotropic grid), a
    three-value list or tuple, or a one- or three-element tensor."""
    kind, items = _flatten(value)
    if kind == "scalar" and _is_real_scalar(value):
        items = items * 3
    elif kind == "tensor":
        if len(items) not in (1, 3):
            raise ValueError("dx tensor must contain either one or three values.")
        items = items * 3 if len(items) == 1 else items
    elif kind == "sequence":
        if len(items) != 3:
            raise ValueError("dx list or tuple must contain exactly three values.")
    else:
        raise TypeError(_SPACING_TYPE)
    spacing = []
    for item in items:
        try:
            spacing.append(float(item))
        except (TypeError, ValueError) as exc:
            raise TypeError(_SPACING_VALUE) from exc
    if not all(math.isfinite(v) and v > 0.0 for v in spacing):
        raise ValueError(_SPACING_VALUE)
    return tuple(spacing)  # type: ignore[return-value]


def _pml_faces(items: List[Any]) -> List[Any]:
    """Six face values [x0, xm, y0, ym, z0, zm] from 4 or 6 listed values."""
    rules: Dict[int, Callable[[List[Any]], List[Any]]] = {4: lambda v: v + [0, 0], 6: lambda v: v}
    if len(items) not in rules:
        raise ValueError(f"Unsupported PML length: {len(items)}. Must be 4 or 6.")
    return rules[len(items)](items)


def _pml_value(v: Any) -> int:
    if isinstance(v, bool):
        raise TypeError(_PML_BOOL)
    try:
        f = float(v)
    except (TypeError, ValueError, OverflowError) as exc:
        raise TypeError(_PML_INT) from exc
    if not (math.isfinite(f) and f.is_integer()):
        raise ValueError(_PML_INT)
    if not 0 <= f <= INT32_MAX:
        raise ValueError(_PML_RANGE)
    return int(f)


def pmlthick_revert(p: Any, er: torch.Tensor) -> torch.Tensor:
    """Six face thicknesses ``[x0, xm, y0, ym, z0, zm]`` as an int32 CPU tensor.

    An integer applies to every active face (the z faces are 0 for a 2D model,
    ``er.shape[2] == 1``); four values are ``[x0, xm, y0, ym]``; six values are
    used as given; a tensor must be one-dimensional with four or six values.
    """
    if isinstance(p, bool):
        raise TypeError(_PML_BOOL)
    two_d = er.shape[2] == 1
    if isinstance(p, numbers.Integral):
        faces = [int(p)] * 4 + [0 if two_d else int(p)] * 2
    elif isinstance(p, torch.Tensor) and p.ndim != 1:
        raise ValueError("PML thickness tensor must be one-dimensional.")
    elif isinstance(p, (list, tuple, torch.Tensor)):
        faces = _pml_faces(_flatten(p)[1])
    else:
        raise TypeError(f"Unsupported PML thickness type: {type(p)}")
    return torch.tensor([_pml_value(v) for v in faces], dtype=torch.int32)


__all__ = ["GridSpacing", "GridSpacingLike", "normalize_grid_spacing", "pmlthick_revert"]
