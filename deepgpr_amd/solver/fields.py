"""Field state allocation and checkpoints (src/DeepGPR/solver/fields.py:15-143)."""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import torch

from ..preprocessing.model_setup import initialization
from .pml import build_pml_phi, pml_face_descriptors


def create_or_separate(fields: Optional[Sequence[torch.Tensor]], nx: int, ny: int, nz: int, nstep: int,
                       device, dtype) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Zero ``(nstep, nx+1, ny+1, nz+1)`` components, or validate three given ones."""
    shape = (nstep, nx + 1, ny + 1, nz + 1)
    if fields is None:
        return tuple(torch.zeros(shape, device=device, dtype=dtype) for _ in range(3))  # type: ignore
    if not isinstance(fields, (list, tuple)) or len(fields) != 3:
        raise ValueError("E and H must each contain exactly three field tensors.")
    for f in fields:
        if not isinstance(f, torch.Tensor) or f.shape != shape:
            raise ValueError(f"Field shape mismatch: got {getattr(f, 'shape', None)}, expected {shape}.")
    return tuple(f.to(device=device, dtype=dtype).contiguous() for f in fields)  # type: ignore


def checkpoint_initial_field(device=None, per_nstep: Optional[int] = None, dx=None, dt=None,
                             source_amplitudes=None, source_location=None, receiver_location=None,
                             er=None, se=None, mr=None, pmlthick=10, fdtd_order: int = 2):
    """Zero E, H and 24 CPML states on the full solver grid for segmented runs."""
    prepared = initialization(device, er, se, mr, source_amplitudes, source_location, receiver_location,
                              dx, dt, pmlthick, fdtd_order)
    nx, ny, nz, nstep = prepared.nx, prepared.ny, prepared.nz, prepared.nstep
    if per_nstep is not None:
        if isinstance(per_nstep, bool) or not isinstance(per_nstep, int):
            raise TypeError("per_nstep must be an integer shot count or None.")
        if not 1 <= per_nstep <= nstep:
            raise ValueError("per_nstep must be between 1 and the input shot count.")
    e = create_or_separate(None, nx, ny, nz, nstep, device, prepared.dtype)
    h = create_or_separate(None, nx, ny, nz, nstep, device, prepared.dtype)
    phi = build_pml_phi(*pml_face_descriptors(nx, ny, nz, prepared.pmlthick), nstep, None, device)
    groups = (e, h, phi)
    if per_nstep is None:
        return groups
    return tuple(tuple(t[:per_nstep] if t.numel() else t for t in g) for g in groups)


def zero_field(*tensors):
    return tuple(torch.zeros_like(t) if t is not None else None for t in tensors)
