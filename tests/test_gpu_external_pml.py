"""compute() on physical models extended into the CPML (restates the
reference's tests/test_external_pml.py:20-31, 118-143 and 193-226 on the HIP
device): coordinate and state validation, disabled faces with a shot subset,
and the finite-difference derivative at a model edge cell, whose material is
replicated into the CPML on every forward."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _problem(dev, shape=(8, 9), pml=(2, 3, 1, 2)):
    model = torch.linspace(3.0, 5.0, steps=int(torch.tensor(shape).prod()), device=dev).reshape(shape)
    sigma = torch.full_like(model, 2.e-4)
    mu = torch.linspace(1., 1.2, steps=model.numel(), device=dev).reshape(shape)
    nt = 35
    source = torch.zeros((1, nt, 1), device=dev)
    source[0, 1:4, 0] = torch.tensor([1., -.5, .25], device=dev)
    z = 0 if len(shape) == 2 or shape[2] == 1 else 2
    args = dict(device=dev, dx=(.02, .017, .015), dt=1.e-11, source_amplitudes=source,
                source_location=torch.tensor([[[0, 3, z]], [[3, 4, z]]], device=dev),
                receiver_location=torch.tensor([[[0, 3, z], [3, 4, z]], [[3, 4, z], [2, 3, z]]], device=dev),
                pmlthick=pml, mu_r=mu, mode=2 if z == 0 else 3)
    return model, sigma, args


def test_physical_coordinate_and_state_validation(cuda_device):
    import deepgpr_amd as dg
    print("[gpu external pml] validation", flush=True)
    model, sigma, args = _problem(cuda_device)
    r = dg.compute(eps_r=model, sigma=sigma, **{**args, "pmlthick": 12})
    assert tuple(r[1][0].shape) == (2, 33, 34, 2)
    shifted = args["source_location"] + torch.tensor([8, 0, 0], device=cuda_device)
    with pytest.raises(ValueError, match="out of range"):
        dg.compute(eps_r=model, sigma=sigma, **{**args, "source_location": shifted})
    with pytest.raises(ValueError, match="z boundaries"):
        dg.compute(eps_r=model, sigma=sigma, **{**args, "pmlthick": [1] * 6})
    with pytest.raises(ValueError, match="Field shape mismatch"):
        dg.compute(eps_r=model, sigma=sigma, E=r[1], **args)
    r = dg.compute(eps_r=model, sigma=sigma, **{**args, "pmlthick": 0, "save_wavefield_history": False})
    assert tuple(r[1][0].shape) == (2, 9, 10, 2)
    assert r[0].numel() == 0


def test_absorption_speed_must_be_fixed_positive_and_finite(cuda_device):
    import deepgpr_amd as dg
    model, sigma, args = _problem(cuda_device)
    for speed in (0., -1., float("nan"), float("inf"), True, torch.tensor(2.e8, requires_grad=True)):
        with pytest.raises((ValueError, TypeError)):
            dg.compute(eps_r=model, sigma=sigma, pml_max_velocity=speed, **args)


@pytest.mark.parametrize("shape,pml", [((8, 9), (0, 2, 1, 0)), ((6, 7, 5), (0, 1, 0, 2, 0, 1)), ((8, 9), 0)])
def test_checkpoint_shot_subset_with_disabled_faces(cuda_device, shape, pml):
    import deepgpr_amd as dg
    print(f"[gpu external pml] shot subset {shape} {pml}", flush=True)
    model, sigma, args = _problem(cuda_device, shape, pml)
    initial = {k: v for k, v in args.items() if k not in ("mode", "mu_r")}
    E, H, PML = dg.checkpoint_initial_field(er=model, se=sigma, mr=args["mu_r"], per_nstep=1, **initial)
    result = dg.compute(eps_r=model, sigma=sigma, E=E, H=H, PML=PML,
                        **{**args, "source_location": args["source_location"][:1],
                           "receiver_location": args["receiver_location"][:1]})
    assert result[-1].shape[0] == 1
    with pytest.raises(ValueError, match="per_nstep"):
        dg.checkpoint_initial_field(er=model, se=sigma, per_nstep=3, **initial)


def test_model_edge_derivative_matches_finite_differences(cuda_device):
    import deepgpr_amd as dg
    print("[gpu external pml] edge-cell finite differences", flush=True)
    model, sigma, args = _problem(cuda_device)
    sigma.fill_(.01)
    model.requires_grad_()
    sigma.requires_grad_()
    result = dg.compute(eps_r=model, sigma=sigma, **args)
    scale = result[-1].detach().double().abs().max()
    (.5 * (result[-1].double() / scale).square().sum()).backward()
    for name, values, steps in [("eps", model, (.04, .02, .01, .005)),
                                ("sigma", sigma, (.003, .0015, .00075, .000375))]:
        errors = []
        for h in steps:
            losses = []
            for sign in (1, -1):
                er, se = model.detach().clone(), sigma.detach().clone()
                (er if name == "eps" else se)[0, 3] += sign * h
                with torch.no_grad():
                    data = dg.compute(eps_r=er, sigma=se, save_wavefield_history=False, **args)[-1]
                losses.append(.5 * (data.double() / scale).square().sum())
            fd = float((losses[0] - losses[1]) / (2 * h))
            ad = float(values.grad[0, 3])
            errors.append(abs(fd - ad) / max(abs(fd), abs(ad), 1.e-20))
        # the reference's criterion: two consecutive steps below 0.3 %
        assert any(max(a, b) < .003 for a, b in zip(errors[:-1], errors[1:])), (name, errors)
