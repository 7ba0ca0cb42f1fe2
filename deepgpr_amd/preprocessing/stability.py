"""CFL limit of the staggered scheme (src/DeepGPR/preprocessing/stability.py:20-114)."""
from __future__ import annotations

import math
from typing import Optional

from ..config.constants import FDTD_STENCIL_COEFFICIENT_SUMS, SPEED_OF_LIGHT, SUPPORTED_FDTD_ORDERS
from ..utils.exceptions import CFLConditionError
from .grid import normalize_grid_spacing


def max_stable_time_step(dx, nx: int, ny: int, nz: int, *, fdtd_order: int = 2,
                         min_er_mr: Optional[float] = None) -> float:
    """``sqrt(min(er mr)) / (c S sqrt(sum over axes with >1 cell of 1/d^2))``."""
    spacing = normalize_grid_spacing(dx)
    if fdtd_order not in SUPPORTED_FDTD_ORDERS:
        raise ValueError("fdtd_order must be one of 2, 4, or 8.")
    inv2 = sum(1.0 / (d * d) for n, d in zip((nx, ny, nz), spacing) if n > 1)
    if inv2 == 0.0:
        raise ValueError("At least one model dimension must contain more than one cell.")
    factor = 1.0
    if min_er_mr is not None:
        v = float(min_er_mr)
        if not math.isfinite(v) or v <= 0.0:
            raise ValueError("epsilon_r * mu_r must be finite and positive for the CFL check.")
        factor = math.sqrt(v)
    return factor / (SPEED_OF_LIGHT * FDTD_STENCIL_COEFFICIENT_SUMS[fdtd_order] * math.sqrt(inv2))


def check_cfl(dx, dt, nx: int, ny: int, nz: int, er=None, mr=None, fdtd_order: int = 2,
              _material_min_er_mr: Optional[float] = None) -> None:
    """Raise :class:`CFLConditionError` when ``dt`` exceeds the stable limit."""
    normalize_grid_spacing(dx)
    step = float(dt)
    if not math.isfinite(step) or step <= 0.0:
        raise ValueError("dt must be a finite positive number.")
    if fdtd_order not in SUPPORTED_FDTD_ORDERS:
        raise ValueError("fdtd_order must be one of 2, 4, or 8.")
    m = None
    if _material_min_er_mr is not None:
        m = float(_material_min_er_mr)
    elif er is not None and mr is not None:
        m = float((er.detach() * mr.detach()).amin().item())
    limit = max_stable_time_step(dx, nx, ny, nz, fdtd_order=fdtd_order, min_er_mr=m)
    if step > limit:
        raise CFLConditionError(f"Does not meet CFL conditions: dt={step:.3e} > dt_max={limit:.3e}")
