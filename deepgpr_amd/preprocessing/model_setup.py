"""Validation of the physical model and its extension into the CPML
(src/DeepGPR/preprocessing/model_setup.py:36-406).

The caller supplies only the physical model; every solve replicates the edge
materials outward into the CPML and pads the one-cell Yee halo. The adjoint of
the replication folds the extended-grid sensitivities back onto the physical
edges, faces and corners, one axis at a time.
"""
from __future__ import annotations

from typing import NamedTuple, Sequence, Tuple

import torch
import torch.nn.functional as F

from ..config.constants import (
    CUDA_MAX_SHOTS_PER_CALL, INT32_MAX, MIN_RELATIVE_PERMITTIVITY, PEC_CONDUCTIVITY_THRESHOLD,
)
from ..utils.logger import get_logger
from .grid import normalize_grid_spacing, pmlthick_revert
from .stability import check_cfl

_LOG = get_logger(__name__)

NATIVE_FLOAT_DTYPE = torch.float32
_HALO = (0, 1, 0, 1, 0, 1)


class PreparedModel(NamedTuple):
    er: torch.Tensor
    se: torch.Tensor
    nx: int
    ny: int
    nz: int
    nt: int
    nstep: int
    nsr: int
    nrx: int
    ere: torch.Tensor
    see: torch.Tensor
    mr: torch.Tensor
    mode: int
    dtype: torch.dtype
    pmlthick: torch.Tensor
    source_amplitudes: torch.Tensor


class ExtendModel(torch.autograd.Function):
    """Replicate-pad a material into the CPML; backward sums the replicas onto
    their owner cells axis by axis (no scatter, deterministic)."""

    @staticmethod
    def forward(ctx, model: torch.Tensor, pml: Sequence[int]) -> torch.Tensor:  # type: ignore[override]
        ctx.shape = tuple(model.shape)
        ctx.pml = tuple(int(p) for p in pml)
        p = ctx.pml
        return F.pad(model[None, None], (p[4], p[5], p[2], p[3], p[0], p[1]), mode="replicate")[0, 0]

    @staticmethod
    def backward(ctx, grad: torch.Tensor):  # type: ignore[override]
        g = grad
        for axis, size in enumerate(ctx.shape):
            lo, hi = ctx.pml[2 * axis], ctx.pml[2 * axis + 1]
            core = g.narrow(axis, lo, size).clone()
            if lo:
                core.select(axis, 0).add_(g.narrow(axis, 0, lo).sum(dim=axis))
            if hi:
                core.select(axis, size - 1).add_(g.narrow(axis, lo + size, hi).sum(dim=axis))
            g = core
        return g.contiguous(), None


def shift_locations(locations: torch.Tensor, pmlthick: torch.Tensor, device) -> torch.Tensor:
    """Physical-model coordinates -> extended-grid coordinates (+ low-face PML)."""
    offset = pmlthick[::2].to(device=device, dtype=torch.int32)
    return (locations.to(device=device, dtype=torch.int32) + offset).contiguous()


def _as3d(t: torch.Tensor, what: str) -> torch.Tensor:
    if t.ndim == 2:
        return t.reshape(*t.shape, 1)
    if t.ndim != 3:
        raise ValueError(f"The shape of {what} should be 2-d or 3-d.")
    return t


def _check_locations(src: torch.Tensor, rx: torch.Tensor, device: torch.device) -> Tuple[int, int, int]:
    if not torch.is_tensor(src) or not torch.is_tensor(rx):
        raise TypeError("source_location and receiver_location must be PyTorch tensors.")
    if src.ndim != 3 or rx.ndim != 3:
        raise ValueError("source_location and receiver_location must have shape (nstep, count, 3).")
    if src.shape[2] != 3 or rx.shape[2] != 3:
        raise ValueError("The last dimension of source_location and receiver_location must be 3.")
    if src.shape[0] != rx.shape[0]:
        raise ValueError("The first dimension (nstep) of source_location and receiver_location should be the same.")
    nstep, nsr, nrx = src.shape[0], src.shape[1], rx.shape[1]
    if nstep < 1:
        raise ValueError("At least one shot is required.")
    if nsr < 1:
        raise ValueError("At least one source per shot is required.")
    if nrx < 1:
        raise ValueError("At least one receiver per shot is required.")
    if device.type == "cuda" and nstep > CUDA_MAX_SHOTS_PER_CALL:
        raise ValueError(f"At most {CUDA_MAX_SHOTS_PER_CALL} shots fit in one compute call; split the batch.")
    for name, t in (("source_location", src), ("receiver_location", rx)):
        if t.dtype == torch.bool or t.is_complex():
            raise TypeError(f"{name} must contain real integer-valued coordinates.")
    return nstep, nsr, nrx


def _waveforms(w: torch.Tensor, nsr: int, device, dtype) -> torch.Tensor:
    if not torch.is_tensor(w):
        raise TypeError("source_amplitudes must be a PyTorch tensor.")
    if w.ndim not in (2, 3):
        raise ValueError("source_amplitudes must have shape (num_waveforms, nt) or (num_waveforms, nt, 1).")
    if w.ndim == 2:
        w = w.unsqueeze(-1)
    if w.shape[2] != 1:
        raise ValueError("The last dimension of source_amplitudes must be 1.")
    if w.shape[0] < 1 or w.shape[1] < 1:
        raise ValueError("source_amplitudes must contain at least one waveform and one time sample.")
    w = w.to(device=device, dtype=dtype).contiguous()
    n = w.shape[0]
    if n > nsr or 1 < n < nsr:
        raise ValueError("The number of source waveforms is incorrect.")
    if n == 1 and nsr != 1:
        w = w.repeat(nsr, 1, 1).contiguous()
    return w


def _integral(t: torch.Tensor) -> torch.Tensor:
    if t.is_floating_point():
        return (torch.isfinite(t) & (t == t.trunc())).all()
    return torch.ones((), dtype=torch.bool, device=t.device)


def initialization(device, er, se, mr, source_amplitudes, source_location, receiver_location,
                   dx, dt, pmlthick, fdtd_order: int = 2) -> PreparedModel:
    """Validate the physical-frame inputs, check CFL on the extended grid and
    extend the materials into the CPML plus the Yee halo."""
    device = torch.device("cpu" if device is None else device)
    dtype = NATIVE_FLOAT_DTYPE
    spacing = normalize_grid_spacing(dx)
    step = float(dt)
    if not (step > 0.0 and step < float("inf")):
        raise ValueError("dt must be a finite positive number.")
    if not torch.is_tensor(er) or not torch.is_tensor(se):
        raise TypeError("er and se must be PyTorch tensors.")
    er = _as3d(er, "epsilon")
    se = _as3d(se, "sigma")
    if mr is not None and not torch.is_tensor(mr):
        raise TypeError("mr must be a PyTorch tensor or None.")
    if mr is not None:
        mr = _as3d(mr, "mr")
    if er.shape != se.shape:
        raise ValueError("The shape of epsilon and sigma should be the same.")
    if any(n < 1 for n in er.shape):
        raise ValueError("The material model dimensions must all be non-empty.")
    nx, ny, nz = er.shape
    mode = 2 if nz == 1 else 3
    er = er.to(device=device, dtype=dtype)
    se = se.to(device=device, dtype=dtype)
    if mr is None:
        mr = torch.ones_like(er, device=device)
    elif mr.shape == er.shape:
        mr = mr.to(device=device, dtype=dtype)
    else:
        raise ValueError("The shape of mr should be the same as epsilon and sigma.")

    nstep, nsr, nrx = _check_locations(source_location, receiver_location, device)
    source_location = source_location.to(device=device).contiguous()
    receiver_location = receiver_location.to(device=device).contiguous()
    source_amplitudes = _waveforms(source_amplitudes, nsr, device, dtype)
    nt = source_amplitudes.shape[1]

    pmlthick = pmlthick_revert(pmlthick, er)
    pml = [int(v) for v in pmlthick.tolist()]
    if nz == 1 and any(pml[4:]):
        raise ValueError("2D models require zero PML thickness on the z boundaries.")
    ext = tuple(n + pml[2 * a] + pml[2 * a + 1] for a, n in enumerate((nx, ny, nz)))
    if any(n >= INT32_MAX for n in ext):
        raise ValueError("Extended model dimensions must fit in positive int32 with a field halo.")

    # one device->host transfer for every scalar the validation needs
    shape = torch.tensor((nx, ny, nz), dtype=torch.int32, device=device)
    stats = torch.stack((
        torch.isfinite(er).all().to(dtype), torch.isfinite(se).all().to(dtype),
        torch.isfinite(mr).all().to(dtype), torch.isfinite(source_amplitudes).all().to(dtype),
        er.amin(), se.amin(), mr.amin(), (er.detach() * mr.detach()).amin(),
        ((source_location >= 0) & (source_location < shape)).all().to(dtype),
        ((receiver_location >= 0) & (receiver_location < shape)).all().to(dtype),
        _integral(source_location).to(dtype), _integral(receiver_location).to(dtype),
        se.amax(),
    )).detach().cpu().tolist()
    for name, ok in zip(("er", "se", "mr", "source_amplitudes"), stats[:4]):
        if not ok:
            raise ValueError(f"`{name}` contains NaN or Inf values.")
    if stats[4] < MIN_RELATIVE_PERMITTIVITY:
        raise ValueError("The values of epsilon is incorrect.(should be greater than 1)")
    if stats[5] < 0:
        raise ValueError("The values of sigma is incorrect.(should be non-negative)")
    if stats[6] <= 0:
        raise ValueError("The values of mr are incorrect (must be positive).")
    if not stats[10]:
        raise ValueError("source_location must contain finite integer-valued coordinates.")
    if not stats[11]:
        raise ValueError("receiver_location must contain finite integer-valued coordinates.")
    if not stats[8]:
        raise ValueError(f"Error: Source coordinates out of range! Valid ranges are x∈[0,{nx}), "
                         f"y∈[0,{ny}), z∈[0,{nz})")
    if not stats[9]:
        raise ValueError(f"Error: Receiver coordinates out of range! Valid ranges are x∈[0,{nx}), "
                         f"y∈[0,{ny}), z∈[0,{nz})")
    if stats[12] > PEC_CONDUCTIVITY_THRESHOLD:
        _LOG.warning("Conductivity above %g S/m is treated as a perfect electric conductor.",
                     PEC_CONDUCTIVITY_THRESHOLD)
    check_cfl(spacing, dt, *ext, fdtd_order=fdtd_order, _material_min_er_mr=stats[7])
    if any(pml):
        er = ExtendModel.apply(er, pml)
        se = ExtendModel.apply(se, pml)
        mr = ExtendModel.apply(mr, pml)
    nx, ny, nz = ext
    return PreparedModel(er, se, nx, ny, nz, nt, nstep, nsr, nrx, F.pad(er, _HALO), F.pad(se, _HALO),
                         F.pad(mr, _HALO), mode, dtype, pmlthick, source_amplitudes)


def extended_shape_of(shape: Sequence[int], pml: Sequence[int]) -> Tuple[int, int, int]:
    return tuple(n + int(pml[2 * a]) + int(pml[2 * a + 1]) for a, n in enumerate(shape))  # type: ignore


def physical_shape_of(shape: Sequence[int], pml: Sequence[int]) -> Tuple[int, int, int]:
    return tuple(n - int(pml[2 * a]) - int(pml[2 * a + 1]) for a, n in enumerate(shape))  # type: ignore
