"""Source wavelets used by the benchmarks and examples
(src/DeepGPR/preprocessing/wavelets.py:46-170)."""
from __future__ import annotations

import math
from typing import Optional

import torch


def _axis(freq, length, dt, peak_time, dtype, device):
    if length < 1:
        raise ValueError("length must be positive")
    dtype = torch.get_default_dtype() if dtype is None else dtype
    t = torch.arange(int(length), dtype=dtype, device=device)
    return t * float(dt) - float(peak_time), float(freq)


def ricker(freq: float, length: int, dt: float, peak_time: float,
           dtype: Optional[torch.dtype] = None, device=None) -> torch.Tensor:
    """``(1 - 2 (pi f t)^2) exp(-(pi f t)^2)`` sampled at ``t = k dt - peak_time``."""
    t, f = _axis(freq, length, dt, peak_time, dtype, device)
    a2 = (math.pi * f * t).square()
    return (1.0 - 2.0 * a2) * torch.exp(-a2)


def gaussian(freq: float, length: int, dt: float, peak_time: float,
             dtype: Optional[torch.dtype] = None, device=None) -> torch.Tensor:
    t, f = _axis(freq, length, dt, peak_time, dtype, device)
    return torch.exp(-(math.pi * f * t).square())


def gaussian_derivative(freq: float, length: int, dt: float, peak_time: float,
                        dtype: Optional[torch.dtype] = None, device=None) -> torch.Tensor:
    """First derivative of the Gaussian, unit peak magnitude."""
    t, f = _axis(freq, length, dt, peak_time, dtype, device)
    a = math.pi * f * t
    w = -2.0 * a * torch.exp(-a.square())
    return w / math.sqrt(2.0 / math.e)
