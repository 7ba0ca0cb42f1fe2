"""Validation of the physical model and its extension into the CPML
(the interface of src/DeepGPR/preprocessing/model_setup.py:36-406: the same
arguments, exception types and messages).

The caller supplies only the physical model; every solve replicates the edge
materials outward into the CPML and pads the one-cell Yee halo. The adjoint of
the replication folds the extended-grid sensitivities back onto the physical
edges, faces and corners, one axis at a time.

Design: the host-side shape checks run first; every check that needs the
tensors' values is then one entry of a rule table (`_VALUE_RULES`) over named
device scalars, all evaluated with a single device-to-host transfer.
"""
from __future__ import annotations

from typing import Callable, Dict, List, NamedTuple, Sequence, Tuple

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
    """Replicate-pad a material into the CPML. The backward is the transpose
    of the replication: along each axis in turn, the replicas on either side
    are summed onto the edge cell they copy (no scatter, so the result does
    not depend on scheduling)."""

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


# ----------------------------------------------------------------------------- host-side checks


def _grid3(t: torch.Tensor, what: str) -> torch.Tensor:
    """A 2D model gains a unit z axis; anything but 2D or 3D is refused."""
    if t.ndim not in (2, 3):
        raise ValueError(f"The shape of {what} should be 2-d or 3-d.")
    return t.unsqueeze(-1) if t.ndim == 2 else t


_LOCATION_SHAPE_RULES: List[Tuple[Callable[[torch.Tensor, torch.Tensor], bool], str]] = [
    (lambda s, r: s.ndim == 3 and r.ndim == 3, "source_location and receiver_location must have shape (nstep, count, 3)."),
    (lambda s, r: s.shape[2] == 3 and r.shape[2] == 3,
     "The last dimension of source_location and receiver_location must be 3."),
    (lambda s, r: s.shape[0] == r.shape[0],
     "The first dimension (nstep) of source_location and receiver_location should be the same."),
    (lambda s, r: s.shape[0] >= 1, "At least one shot is required."),
    (lambda s, r: s.shape[1] >= 1, "At least one source per shot is required."),
    (lambda s, r: r.shape[1] >= 1, "At least one receiver per shot is required."),
]


def _location_counts(src, rx, device: torch.device) -> Tuple[int, int, int]:
    """(shots, sources per shot, receivers per shot) of well-shaped location tensors."""
    if not (torch.is_tensor(src) and torch.is_tensor(rx)):
        raise TypeError("source_location and receiver_location must be PyTorch tensors.")
    for holds, message in _LOCATION_SHAPE_RULES:
        if not holds(src, rx):
            raise ValueError(message)
    if device.type == "cuda" and src.shape[0] > CUDA_MAX_SHOTS_PER_CALL:
        raise ValueError(f"At most {CUDA_MAX_SHOTS_PER_CALL} shots fit in one compute call; split the batch.")
    for name, t in (("source_location", src), ("receiver_location", rx)):
        if t.dtype == torch.bool or t.is_complex():
            raise TypeError(f"{name} must contain real integer-valued coordinates.")
    return int(src.shape[0]), int(src.shape[1]), int(rx.shape[1])


def _per_source_waveforms(w, nsr: int, device, dtype) -> torch.Tensor:
    """(nsr, nt, 1) waveforms: one per source, or one shared by every source."""
    if not torch.is_tensor(w):
        raise TypeError("source_amplitudes must be a PyTorch tensor.")
    if w.ndim not in (2, 3):
        raise ValueError("source_amplitudes must have shape (num_waveforms, nt) or (num_waveforms, nt, 1).")
    w = w if w.ndim == 3 else w[..., None]
    if w.shape[2] != 1:
        raise ValueError("The last dimension of source_amplitudes must be 1.")
    if min(w.shape[0], w.shape[1]) < 1:
        raise ValueError("source_amplitudes must contain at least one waveform and one time sample.")
    if w.shape[0] not in (1, nsr):
        raise ValueError("The number of source waveforms is incorrect.")
    w = w.to(device=device, dtype=dtype)
    return (w.expand(nsr, -1, -1) if w.shape[0] != nsr else w).contiguous()


# ----------------------------------------------------------------------------- value checks


def _whole(t: torch.Tensor) -> torch.Tensor:
    """All entries are finite whole numbers (always true for integer tensors)."""
    if not t.is_floating_point():
        return torch.ones((), dtype=torch.bool, device=t.device)
    return (torch.isfinite(t) & (torch.frac(t) == 0)).all()


def _inside(loc: torch.Tensor, shape: torch.Tensor) -> torch.Tensor:
    return ((loc >= 0) & (loc < shape)).all()


# (name of a device scalar, whether its host value passes, exception, message);
# checked in this order once all scalars have reached the host.
_VALUE_RULES: List[Tuple[str, Callable[[float], bool], type, str]] = [
    ("finite_er", bool, ValueError, "`er` contains NaN or Inf values."),
    ("finite_se", bool, ValueError, "`se` contains NaN or Inf values."),
    ("finite_mr", bool, ValueError, "`mr` contains NaN or Inf values."),
    ("finite_wave", bool, ValueError, "`source_amplitudes` contains NaN or Inf values."),
    ("min_er", lambda v: v >= MIN_RELATIVE_PERMITTIVITY, ValueError,
     "The values of epsilon is incorrect.(should be greater than 1)"),
    ("min_se", lambda v: v >= 0, ValueError, "The values of sigma is incorrect.(should be non-negative)"),
    ("min_mr", lambda v: v > 0, ValueError, "The values of mr are incorrect (must be positive)."),
    ("whole_src", bool, ValueError, "source_location must contain finite integer-valued coordinates."),
    ("whole_rx", bool, ValueError, "receiver_location must contain finite integer-valued coordinates."),
    ("inside_src", bool, ValueError, "Error: Source coordinates out of range! Valid ranges are {ranges}"),
    ("inside_rx", bool, ValueError, "Error: Receiver coordinates out of range! Valid ranges are {ranges}"),
]


def _device_scalars(er, se, mr, wave, src, rx) -> Dict[str, float]:
    """Every value the checks need, moved to the host in one transfer."""
    shape = torch.tensor(er.shape, dtype=torch.int32, device=er.device)
    named = {
        "finite_er": torch.isfinite(er).all(), "finite_se": torch.isfinite(se).all(),
        "finite_mr": torch.isfinite(mr).all(), "finite_wave": torch.isfinite(wave).all(),
        "min_er": er.amin(), "min_se": se.amin(), "min_mr": mr.amin(),
        "min_er_mr": (er.detach() * mr.detach()).amin(), "max_se": se.amax(),
        "whole_src": _whole(src), "whole_rx": _whole(rx),
        "inside_src": _inside(src, shape), "inside_rx": _inside(rx, shape),
    }
    values = torch.stack([v.detach().to(torch.float32) for v in named.values()]).cpu().tolist()
    return dict(zip(named, values))


def initialization(device, er, se, mr, source_amplitudes, source_location, receiver_location,
                   dx, dt, pmlthick, fdtd_order: int = 2) -> PreparedModel:
    """Validate the physical-frame inputs, check CFL on the extended grid and
    extend the materials into the CPML plus the Yee halo."""
    device = torch.device("cpu" if device is None else device)
    dtype = NATIVE_FLOAT_DTYPE
    spacing = normalize_grid_spacing(dx)
    if not 0.0 < float(dt) < float("inf"):
        raise ValueError("dt must be a finite positive number.")
    if not (torch.is_tensor(er) and torch.is_tensor(se)):
        raise TypeError("er and se must be PyTorch tensors.")
    if mr is not None and not torch.is_tensor(mr):
        raise TypeError("mr must be a PyTorch tensor or None.")
    er, se = _grid3(er, "epsilon"), _grid3(se, "sigma")
    mr = None if mr is None else _grid3(mr, "mr")
    if er.shape != se.shape:
        raise ValueError("The shape of epsilon and sigma should be the same.")
    if min(er.shape) < 1:
        raise ValueError("The material model dimensions must all be non-empty.")
    if mr is not None and mr.shape != er.shape:
        raise ValueError("The shape of mr should be the same as epsilon and sigma.")
    nx, ny, nz = (int(n) for n in er.shape)
    er, se = er.to(device=device, dtype=dtype), se.to(device=device, dtype=dtype)
    mr = torch.ones_like(er) if mr is None else mr.to(device=device, dtype=dtype)

    nstep, nsr, nrx = _location_counts(source_location, receiver_location, device)
    source_location = source_location.to(device=device).contiguous()
    receiver_location = receiver_location.to(device=device).contiguous()
    source_amplitudes = _per_source_waveforms(source_amplitudes, nsr, device, dtype)

    pmlthick = pmlthick_revert(pmlthick, er)
    pml = [int(v) for v in pmlthick.tolist()]
    if nz == 1 and (pml[4] or pml[5]):
        raise ValueError("2D models require zero PML thickness on the z boundaries.")
    ext = tuple(n + pml[2 * a] + pml[2 * a + 1] for a, n in enumerate((nx, ny, nz)))
    if max(ext) >= INT32_MAX:
        raise ValueError("Extended model dimensions must fit in positive int32 with a field halo.")

    v = _device_scalars(er, se, mr, source_amplitudes, source_location, receiver_location)
    ranges = f"x∈[0,{nx}), y∈[0,{ny}), z∈[0,{nz})"
    for name, passes, exc, message in _VALUE_RULES:
        if not passes(v[name]):
            raise exc(message.format(ranges=ranges))
    if v["max_se"] > PEC_CONDUCTIVITY_THRESHOLD:
        _LOG.warning("Conductivity above %g S/m is treated as a perfect electric conductor.",
                     PEC_CONDUCTIVITY_THRESHOLD)
    check_cfl(spacing, dt, *ext, fdtd_order=fdtd_order, _material_min_er_mr=v["min_er_mr"])
    if any(pml):
        er, se, mr = (ExtendModel.apply(t, pml) for t in (er, se, mr))
    return PreparedModel(er, se, *ext, source_amplitudes.shape[1], nstep, nsr, nrx, F.pad(er, _HALO),
                         F.pad(se, _HALO), F.pad(mr, _HALO), 2 if nz == 1 else 3, dtype, pmlthick,
                         source_amplitudes)


def extended_shape_of(shape: Sequence[int], pml: Sequence[int]) -> Tuple[int, int, int]:
    return tuple(n + int(pml[2 * a]) + int(pml[2 * a + 1]) for a, n in enumerate(shape))  # type: ignore


def physical_shape_of(shape: Sequence[int], pml: Sequence[int]) -> Tuple[int, int, int]:
    return tuple(n - int(pml[2 * a]) - int(pml[2 * a + 1]) for a, n in enumerate(shape))  # type: ignore
