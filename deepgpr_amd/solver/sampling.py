"""Temporal sampling interval of the gradient history
(src/DeepGPR/solver/sampling.py:58-134): ``S <= floor(1 / (4 f_max dt))``,
``f_max`` the highest frequency whose source amplitude spectrum still exceeds a
fraction of its peak (zero-padded FFT, linear interpolation between bins)."""
from __future__ import annotations

import math
from typing import Any

import torch

from ..config.constants import SOURCE_SPECTRUM_CUTOFF_FRACTION

AUTO_SAMPLING_INTERVAL = "auto"
_OVERSAMPLING = 4


def _rows(w: Any) -> torch.Tensor:
    w = torch.as_tensor(w).detach()
    if w.ndim == 3:
        if w.shape[-1] != 1:
            raise ValueError("source_amplitudes must have shape (n, nt, 1), (n, nt) or (nt,).")
        w = w[..., 0]
    elif w.ndim == 1:
        w = w.unsqueeze(0)
    elif w.ndim != 2:
        raise ValueError("source_amplitudes must have shape (n, nt, 1), (n, nt) or (nt,).")
    if w.shape[-1] < 2:
        raise ValueError("source_amplitudes needs at least two time samples.")
    return w.to(device="cpu", dtype=torch.float64)


def _dt(dt: Any) -> float:
    v = float(dt)
    if not math.isfinite(v) or v <= 0.0:
        raise ValueError("dt must be a positive finite number.")
    return v


def source_max_frequency(source_amplitudes: Any, dt: Any, *,
                         cutoff_fraction: float = SOURCE_SPECTRUM_CUTOFF_FRACTION) -> float:
    if not 0.0 < float(cutoff_fraction) < 1.0:
        raise ValueError("cutoff_fraction must lie in (0, 1).")
    step = _dt(dt)
    rows = _rows(source_amplitudes)
    if not bool(torch.isfinite(rows).all()):
        raise ValueError("source_amplitudes must be finite.")
    n = rows.shape[-1]
    nfft = 1 << max(0, (_OVERSAMPLING * n - 1).bit_length())
    spec = torch.fft.rfft(rows, n=nfft, dim=-1).abs()
    df = 1.0 / (nfft * step)
    best = 0.0
    for row in spec:
        peak = float(row.max())
        if peak <= 0.0:
            continue
        thr = float(cutoff_fraction) * peak
        last = int(torch.nonzero(row >= thr).flatten()[-1])
        pos = float(last)
        if last + 1 < row.numel():
            hi, lo = float(row[last]), float(row[last + 1])
            if hi > lo:
                pos += (hi - thr) / (hi - lo)
        best = max(best, pos * df)
    return best


def recommended_sampling_interval(source_amplitudes: Any, dt: Any, *,
                                  cutoff_fraction: float = SOURCE_SPECTRUM_CUTOFF_FRACTION) -> int:
    step = _dt(dt)
    f = source_max_frequency(source_amplitudes, step, cutoff_fraction=cutoff_fraction)
    if f <= 0.0:
        return 1
    return max(1, int(math.floor(1.0 / (4.0 * f * step))))
