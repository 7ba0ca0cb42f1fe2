"""Forward-only central differences of the whole physical model with CPML
against the adjoint gradient (restates the reference's
tests/test_pml_total_derivative.py:31-232 on the HIP device, on the 2D Ez-TM
kernels and on the general kernels). Every difference rebuilds the replicated
materials through compute(), with the CPML absorption speed fixed; the
directions include the boundary cells. Receivers over 160 (2D) / 90 (3D)
steps, and one step from random nonzero states with the final E as the
objective. Agreement: relative error below 0.2 % at two adjacent step sizes.
The reference exempts one direction, random and joint on the 2D receiver
case, where float32 rounding of the forward limits the difference; it is
left out here."""
import contextlib
import math

import pytest
import torch

pytestmark = pytest.mark.gpu

STEPS = (3.0, 2.0, 1.0, .5, .25, .125, .0625, .03125, .015625, .0078125)
SPEED_OF_LIGHT = 299792458.0


@contextlib.contextmanager
def _path(path):
    from deepgpr_amd.solver import modeling
    original = modeling._tm2d_applies
    if path == "general":
        modeling._tm2d_applies = lambda *args, **kwargs: False
    try:
        yield
    finally:
        modeling._tm2d_applies = original


def _problem(dev, dims, nonzero_state):
    import deepgpr_amd as dg
    from deepgpr_amd.solver.autograd import TM2D_INACTIVE_STATE_INDICES
    shape = (18, 20) if dims == 2 else (8, 9, 7)
    coords = torch.meshgrid(*[torch.linspace(0, 1, n, device=dev) for n in shape], indexing="ij")
    eps = 4.0 + sum((.3 / (axis + 1)) * c for axis, c in enumerate(coords))
    sigma = .01 + sum((.002 / (axis + 1)) * c for axis, c in enumerate(coords))
    nt = 1 if nonzero_state else (160 if dims == 2 else 90)
    dt = 2.5e-11 if dims == 2 else 1.5e-11
    z = 0 if dims == 2 else 3
    wave = (torch.zeros((1, nt, 1), device=dev) if nonzero_state
            else dg.ricker(4.e8, nt, dt, 4.e-10).to(dev).reshape(1, nt, 1))
    common = dict(device=dev, dx=(.020, .017, .015), dt=dt, source_amplitudes=wave,
                  source_location=torch.tensor([[[1, 3, z]]], device=dev, dtype=torch.int32),
                  receiver_location=torch.tensor([[[0, 4, z], [2, 6, z], [5, 3, z]]], device=dev, dtype=torch.int32),
                  pmlthick=[3, 2, 4, 1] if dims == 2 else [2, 1, 1, 2, 2, 1], fdtd_order=4,
                  mode=2 if dims == 2 else 3, pml_max_velocity=.75 * SPEED_OF_LIGHT)
    states = None
    if nonzero_state:
        initial = {k: v for k, v in common.items() if k not in ("mode", "pml_max_velocity")}
        groups = dg.checkpoint_initial_field(er=eps, se=sigma, **initial)
        gen = torch.Generator().manual_seed(720 + dims)
        states = [(torch.randn(t.shape, dtype=t.dtype, generator=gen) * 1.e-3).to(dev) for g in groups for t in g]
        if dims == 2:
            for index in TM2D_INACTIVE_STATE_INDICES:  # identically zero in the 2D Ez-TM system
                states[index].zero_()
    return eps, sigma, common, states


def _run(eps, sigma, common, states, path, *, rhs="stored", history=True):
    import deepgpr_amd as dg
    options = dict(common, wavefield_rhs_history=rhs, save_wavefield_history=history)
    if states is not None:
        owned = [t.detach().clone() for t in states]
        options.update(E=tuple(owned[:3]), H=tuple(owned[3:6]), PML=tuple(owned[6:]))
    with _path(path):
        return dg.compute(eps_r=eps, sigma=sigma, **options)


def _directions(shape, dev):
    coords = torch.meshgrid(*[torch.linspace(0, 1, n, dtype=torch.float64, device=dev) for n in shape],
                            indexing="ij")
    smooth = .6 + sum(.15 * torch.cos(math.pi * c) for c in coords)
    edge = torch.zeros(shape, dtype=torch.float64, device=dev)
    for axis in range(len(shape)):
        edge.select(axis, 0).fill_(1.)
        edge.select(axis, shape[axis] - 1).fill_(.7)
    gen = torch.Generator().manual_seed(1052 + len(shape))
    random = torch.randn(shape, dtype=torch.float64, generator=gen).to(dev)
    return {name: v / v.abs().max() for name, v in (("edge", edge), ("random", random), ("smooth", smooth))}


def _objective(result, target, scale, nonzero_state):
    if nonzero_state:
        return .5 * sum(t.double().square().sum() for t in result[1])
    return .5 * ((result[-1].double() - target) / scale).square().sum()


@pytest.mark.parametrize("nonzero_state", [False, True], ids=["receivers", "one_step_states"])
@pytest.mark.parametrize("path,dims", [("fast", 2), ("general", 2), ("general", 3)])
@pytest.mark.parametrize("rhs", ["stored", "reconstructed"])
def test_whole_model_derivatives_match_central_differences(cuda_device, path, dims, rhs, nonzero_state):
    print(f"[gpu pml derivative] {path} {dims}D {rhs} states={nonzero_state}", flush=True)
    base_eps, base_sigma, common, states = _problem(cuda_device, dims, nonzero_state)
    with torch.no_grad():
        baseline = _run(base_eps, base_sigma, common, states, path, history=False)
        target = .82 * baseline[-1].double()
        scale = baseline[-1].double().abs().max().clamp_min(1.e-20)
    eps = base_eps.detach().clone().requires_grad_()
    sigma = base_sigma.detach().clone().requires_grad_()
    result = _run(eps, sigma, common, states, path, rhs=rhs)
    if dims == 2:
        assert bool(result[-1].grad_fn.config.tm2d_fast_path) == (path == "fast")
    _objective(result, target, scale, nonzero_state).backward()
    assert bool(torch.isfinite(eps.grad).all()) and bool(torch.isfinite(sigma.grad).all())
    failures = []
    for kind, unit in _directions(base_eps.shape, cuda_device).items():
        for parameter in ("eps", "sigma", "joint"):
            if dims == 2 and not nonzero_state and kind == "random" and parameter == "joint":
                continue
            de = unit * .25 if parameter in ("eps", "joint") else torch.zeros_like(unit)
            ds = unit * .0025 if parameter in ("sigma", "joint") else torch.zeros_like(unit)
            adjoint = float((eps.grad.double() * de).sum() + (sigma.grad.double() * ds).sum())
            errors = []
            for h in STEPS:
                losses = []
                for sign in (1, -1):
                    with torch.no_grad():
                        out = _run((base_eps.double() + sign * h * de).float(),
                                   (base_sigma.double() + sign * h * ds).float(), common, states, path,
                                   history=False)
                    losses.append(float(_objective(out, target, scale, nonzero_state)))
                fd = (losses[0] - losses[1]) / (2 * h)
                errors.append(abs(fd - adjoint) / max(abs(fd), abs(adjoint), 1.e-20))
            if not any(max(a, b) < .002 for a, b in zip(errors[:-1], errors[1:])):
                failures.append((kind, parameter, [f"{e:.2e}" for e in errors]))
    assert not failures, failures
