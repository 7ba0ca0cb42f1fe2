"""CFS-CPML profiles, update coefficients and auxiliary state
(src/DeepGPR/solver/pml.py:42-449).

Profiles follow the complex-frequency-shifted CPML recursion of Roden & Gedney
(2000) in the gprMax convention: alpha = 0, kappa = 1 and a quartic sigma
graded from 0 to sigma_max = 0.8 (m + 1) eps0 v_ref / d. Electric samples sit
at k/p, magnetic samples at (k + 1/2)/p, k = 0..p-1. The coefficients are
computed with the same float32 torch operations, in the same order, as the
reference so that both produce the same bits; the FDTD kernels take them as
inputs.
"""
from __future__ import annotations

import math
import threading
from typing import Any, Dict, List, Optional, Sequence, Tuple

import torch

from ..config.constants import (
    CPML_COEFFICIENTS_PER_CELL, CPML_FACE_COUNT, CPML_PHI_TENSOR_COUNT, CPML_PHI_TENSORS_PER_FACE,
    CPML_POLE_COUNT, CPML_SCALING_PROFILES, CPML_SIGMA_MAX_FACTOR, SPEED_OF_LIGHT,
    VACUUM_PERMEABILITY, VACUUM_PERMITTIVITY,
)
from ..preprocessing.grid import normalize_grid_spacing

e0 = VACUUM_PERMITTIVITY
m0 = VACUUM_PERMEABILITY
ShapeTuple = Tuple[int, ...]


class CFSParameter:
    """One CFS parameter profile: ``scalingprofile`` names the polynomial order."""

    scalingprofiles = CPML_SCALING_PROFILES

    def __init__(self, ID=None, scaling="polynomial", scalingprofile=None, min=0, max=0):
        self.ID, self.scaling, self.scalingprofile, self.min, self.max = ID, scaling, scalingprofile, min, max


class CFS:
    """Single-pole CFS profile generator (alpha const 0, kappa const 1, quartic sigma)."""

    def __init__(self, device):
        self.alpha = CFSParameter(ID="alpha", scalingprofile="constant")
        self.kappa = CFSParameter(ID="kappa", scalingprofile="constant", min=1, max=1)
        self.sigma = CFSParameter(ID="sigma", scalingprofile="quartic", min=0, max=None)
        self.device = device

    def calculate_sigmamax(self, d, er, mr):
        with torch.no_grad():
            m = CFSParameter.scalingprofiles[self.sigma.scalingprofile]
            self.sigma.max = (CPML_SIGMA_MAX_FACTOR * (m + 1)) / (((m0 / e0) ** 0.5) * d * torch.sqrt(er * mr))

    def scaling_polynomial(self, order, Evalues, Hvalues):
        """Staggered samples of ``x ** order``: E at k/p, H at (k+1/2)/p, k = 0..p."""
        p = len(Evalues) - 1
        grid = (torch.linspace(0, p + 0.5, steps=2 * (p + 1)) / p) ** order
        return grid[0:-1:2].to(self.device), grid[1::2].to(self.device)

    def calculate_values(self, thickness, parameter):
        e = torch.zeros(thickness + 1, device=self.device)
        h = torch.zeros(thickness + 1, device=self.device)
        if parameter.scalingprofile == "constant":
            e += parameter.max
            h += parameter.max
        elif parameter.scaling == "polynomial":
            e, h = self.scaling_polynomial(CFSParameter.scalingprofiles[parameter.scalingprofile], e, h)
            lo, hi = parameter.min, parameter.max
            e = e * (hi - lo) + lo
            h = h * (hi - lo) + lo
        return e[:-1], h[:-1]


def calculate_pml_update_coeffs(cfs, R1, R2, aver, avmr, dt, d, thickness):
    """Fill RA, RB, RE, RF of the electric (R1) and magnetic (R2) profiles."""
    if not cfs.sigma.max:
        cfs.calculate_sigmamax(d, aver, avmr)
    ea, ha = cfs.calculate_values(thickness, cfs.alpha)
    ek, hk = cfs.calculate_values(thickness, cfs.kappa)
    es, hs = cfs.calculate_values(thickness, cfs.sigma)
    for R, a, k, s in ((R1, ea, ek, es), (R2, ha, hk, hs)):
        den = (2 * e0 * k) + dt * (a * k + s)
        R[0, 0, :] = (2 * e0 + dt * a) / den
        R[1, 0, :] = (2 * e0 * k) / den
        R[2, 0, :] = ((2 * e0 * k) - dt * (a * k + s)) / den
        R[3, 0, :] = (2 * s * dt) / (k * den)


def _face_box(face: int, p: int, nx: int, ny: int, nz: int) -> ShapeTuple:
    boxes = ((p, 0, p, 0, ny, 0, nz), (p, nx - p, nx, 0, ny, 0, nz),
             (p, 0, nx, 0, p, 0, nz), (p, 0, nx, ny - p, ny, 0, nz),
             (p, 0, nx, 0, ny, 0, p), (p, 0, nx, 0, ny, nz - p, nz))
    return boxes[face]


def pml_face_descriptors(nx: int, ny: int, nz: int, pmlthick: Any) -> Tuple[torch.Tensor, ...]:
    """``(p, x0, x1, y0, y1, z0, z1)`` int32 per face; empty for a disabled face."""
    pml = [int(v) for v in (pmlthick.tolist() if torch.is_tensor(pmlthick) else pmlthick)]
    return tuple(torch.tensor(_face_box(f, pml[f], nx, ny, nz), dtype=torch.int) if pml[f] > 0
                 else torch.empty(0) for f in range(CPML_FACE_COUNT))


_PROFILE_CACHE: Dict[tuple, tuple] = {}
_PROFILE_LOCK = threading.Lock()


def _face_profiles(p: int, dt: float, d: float, v_ref: float, device, dtype):
    """(electric, magnetic) coefficients ``(4, 1, p)`` of one face. They depend
    only on these arguments, so they are computed once (with the same float32
    operations, on the device) and copied for every later call; a copy on
    another stream waits for the computation through an event."""
    dev = torch.device(device)
    key = (p, float(dt), float(d), v_ref, str(dev), dtype)
    with _PROFILE_LOCK:
        hit = _PROFILE_CACHE.get(key)
    if hit is None:
        cfs = CFS(device=dev)
        order = CFSParameter.scalingprofiles[cfs.sigma.scalingprofile]
        sigma_max = CPML_SIGMA_MAX_FACTOR * (order + 1) * e0 * v_ref / d
        finfo = torch.finfo(torch.get_default_dtype())
        if not math.isfinite(sigma_max) or not finfo.tiny <= sigma_max <= finfo.max:
            raise ValueError("max_velocity produces an unrepresentable PML absorption profile.")
        cfs.sigma.max = sigma_max
        shape = (CPML_COEFFICIENTS_PER_CELL, CPML_POLE_COUNT, p)
        el = torch.zeros(shape, device=dev, dtype=dtype)
        ma = torch.zeros(shape, device=dev, dtype=dtype)
        calculate_pml_update_coeffs(cfs, el, ma, 1.0, 1.0, dt, d, p)
        if not bool(torch.isfinite(el).all() & torch.isfinite(ma).all()):
            raise ValueError("max_velocity produces non-finite PML update coefficients.")
        ready = None
        if dev.type == "cuda":
            ready = torch.cuda.Event()
            ready.record(torch.cuda.current_stream(dev))
        hit = (el, ma, ready)
        with _PROFILE_LOCK:
            _PROFILE_CACHE.setdefault(key, hit)
    el, ma, ready = hit
    if ready is not None:
        torch.cuda.current_stream(dev).wait_event(ready)
    return el.clone(), ma.clone()


def build_pml_coeffs(eps_r, mu_r, dt, dx, nx, ny, nz, pmlthick, device, dtype, *,
                     max_velocity=SPEED_OF_LIGHT):
    """Six face descriptors followed by (electric, magnetic) coefficients per
    face, each ``(4, 1, thickness)`` (empty for disabled faces). The profiles
    depend on the fixed reference speed only, not on the materials."""
    spacing = normalize_grid_spacing(dx)
    if isinstance(max_velocity, bool) or getattr(max_velocity, "requires_grad", False):
        raise TypeError("max_velocity must be a fixed finite positive scalar.")
    v_ref = float(max_velocity)
    if not math.isfinite(v_ref) or v_ref <= 0.0:
        raise ValueError("max_velocity must be a finite positive scalar.")
    pml = [int(v) for v in pmlthick.tolist()]
    descriptors: List[torch.Tensor] = []
    coefficients: List[torch.Tensor] = []
    for face in range(CPML_FACE_COUNT):
        p = pml[face]
        if p <= 0:
            descriptors.append(torch.empty(0))
            coefficients.extend((torch.empty(0), torch.empty(0)))
            continue
        descriptors.append(torch.tensor(_face_box(face, p, nx, ny, nz), dtype=torch.int))
        coefficients.extend(_face_profiles(p, dt, spacing[face // 2], v_ref, device, dtype))
    return (*descriptors, *coefficients)


def buildpmlcoeffs(*args, **kwargs):
    """The older name of build_pml_coeffs, which also takes the materials as
    ``er`` / ``mr`` keywords (src/DeepGPR/solver/pml.py:301-307)."""
    for old, new in (("er", "eps_r"), ("mr", "mu_r")):
        if old in kwargs:
            if new in kwargs:
                raise TypeError(f"Specify only one of {new} and its older name {old}.")
            kwargs[new] = kwargs.pop(old)
    return build_pml_coeffs(*args, **kwargs)


def expected_phi_shapes(face: int, descriptor: torch.Tensor, nstep: int) -> Tuple[Optional[ShapeTuple], ...]:
    """Shapes of (E_phi1, E_phi2, H_phi1, H_phi2) of one face (None: disabled)."""
    if descriptor.numel() == 0:
        return (None,) * CPML_PHI_TENSORS_PER_FACE
    a = int(descriptor[2] - descriptor[1])
    b = int(descriptor[4] - descriptor[3])
    c = int(descriptor[6] - descriptor[5])
    axis = face // 2
    if axis == 0:
        return ((nstep, a + 1, b, c + 1), (nstep, a + 1, b + 1, c), (nstep, a, b + 1, c), (nstep, a, b, c + 1))
    if axis == 1:
        return ((nstep, a, b + 1, c + 1), (nstep, a + 1, b + 1, c), (nstep, a + 1, b, c), (nstep, a, b, c + 1))
    return ((nstep, a, b + 1, c + 1), (nstep, a + 1, b, c + 1), (nstep, a + 1, b, c), (nstep, a, b + 1, c))


def build_pml_phi(x0, xm, y0, ym, z0, zm, nstep, PML, device):
    """Allocate zero CPML auxiliaries or validate the 24 supplied ones.

    Supplied float32 tensors already on ``device`` are returned as-is: the
    solver advances them in place."""
    descriptors = (x0, xm, y0, ym, z0, zm)
    if PML is None:
        PML = (None,) * CPML_PHI_TENSOR_COUNT
    elif not isinstance(PML, (list, tuple)) or len(PML) != CPML_PHI_TENSOR_COUNT:
        raise ValueError("PML must contain exactly 24 CPML auxiliary tensors.")
    out = []
    for face, desc in enumerate(descriptors):
        group = PML[4 * face: 4 * face + 4]
        given = [g is not None for g in group]
        if any(given) and not all(given):
            raise ValueError(f"CPML face {face} must provide all four auxiliary tensors or none.")
        if all(given) and not all(torch.is_tensor(g) for g in group):
            raise TypeError(f"CPML face {face} state must contain PyTorch tensors.")
        if all(given) and desc.numel() == 0 and any(g.numel() != 0 for g in group):
            raise ValueError(f"CPML face {face} state must be empty for a zero-thickness boundary.")
        shapes = expected_phi_shapes(face, desc, nstep)
        if group[0] is not None:
            tensors = [g.contiguous() for g in group]
        elif desc.numel():
            tensors = [torch.zeros(s, dtype=torch.float, device=device) for s in shapes]
        else:
            tensors = [torch.empty(0) for _ in range(4)]
        for k, (t, s) in enumerate(zip(tensors, shapes)):
            if s is None:
                if t.numel() != 0:
                    raise ValueError(f"CPML face {face} component {k} must be empty.")
            elif tuple(t.shape) != s:
                raise ValueError(f"CPML face {face} component {k} has shape {tuple(t.shape)}; expected {s}.")
        out.extend(tensors)
    return tuple(t.to(device=device, dtype=torch.float32).contiguous() for t in out)


def pml_phi_element_count(nx: int, ny: int, nz: int, nstep: int, pml: Sequence[int]) -> int:
    """Exact float count of all 24 auxiliary tensors (extended sizes nx, ny, nz)."""
    total = 0
    for face in range(6):
        p = int(pml[face])
        if p <= 0:
            continue
        desc = torch.tensor(_face_box(face, p, nx, ny, nz))
        for s in expected_phi_shapes(face, desc, nstep):
            total += math.prod(s)
    return total
