"""The CPML absorption speed and profile builders (restates the reference's
tests/test_pml_profile_config.py:26-172; no GPU needed): invalid speeds are
refused before any model work, unrepresentable profiles raise, extreme
speeds never give nonfinite coefficients, the profiles do not depend on the
materials, the default is the vacuum speed, and the older builder name takes
the older material keywords."""
import inspect

import pytest
import torch

import deepgpr_amd as dg
from deepgpr_amd.solver.pml import build_pml_coeffs

SPEED_OF_LIGHT = 299792458.0
INVALID = (0.0, -1.0, float("nan"), float("inf"), -float("inf"), True, False, None, "invalid", 1.0 + 2.0j,
           torch.tensor(SPEED_OF_LIGHT, requires_grad=True))


def _arguments():
    eps = torch.full((8, 9, 6), 4.0, dtype=torch.float32)
    return dict(eps_r=eps, mu_r=torch.ones_like(eps), dt=1.0e-11, dx=(0.02, 0.017, 0.015), nx=8, ny=9, nz=6,
                pmlthick=torch.tensor([2, 1, 0, 2, 1, 0], dtype=torch.int32), device="cpu", dtype=torch.float32)


@pytest.fixture
def default_dtype():
    previous = torch.get_default_dtype()
    yield
    torch.set_default_dtype(previous)


def test_compute_rejects_invalid_speeds_before_model_work(monkeypatch):
    from deepgpr_amd.solver import modeling

    def untouched(*args, **kwargs):
        raise AssertionError("initialization ran")

    monkeypatch.setattr(modeling, "initialization", untouched)
    for speed in INVALID:
        with pytest.raises((TypeError, ValueError)):
            dg.compute(device="cpu", pml_max_velocity=speed)


@pytest.mark.parametrize("speed", INVALID, ids=range(len(INVALID)))
def test_builder_rejects_invalid_speeds(speed):
    with pytest.raises((TypeError, ValueError)):
        build_pml_coeffs(**_arguments(), max_velocity=speed)


@pytest.mark.parametrize("speed", [1.0e308, 1.0e-320, 1.0e-250])
def test_unrepresentable_float32_profiles_are_rejected(default_dtype, speed):
    torch.set_default_dtype(torch.float32)
    with pytest.raises(ValueError, match="unrepresentable"):
        build_pml_coeffs(**_arguments(), max_velocity=speed)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_extreme_speeds_never_give_nonfinite_coefficients(default_dtype, dtype):
    torch.set_default_dtype(dtype)
    for speed in (1.0e-320, 1.0e-250, 1.0e30, 1.0e308):
        try:
            arrays = build_pml_coeffs(**_arguments(), max_velocity=speed)
        except ValueError:
            continue
        assert all(bool(torch.isfinite(c).all()) for c in arrays[6:]), speed


def test_profiles_do_not_depend_on_the_materials():
    args = _arguments()
    reference = build_pml_coeffs(**args, max_velocity=2.0e8)
    args["eps_r"] = torch.linspace(1.0, 15.0, args["eps_r"].numel()).reshape(args["eps_r"].shape).requires_grad_()
    args["mu_r"] = torch.full_like(args["mu_r"], 3.0, requires_grad=True)
    changed = build_pml_coeffs(**args, max_velocity=2.0e8)
    for before, after in zip(reference, changed):
        assert torch.equal(before, after) and not after.requires_grad
    slower = build_pml_coeffs(**args, max_velocity=1.0e8)
    assert any(not torch.equal(a, b) for a, b in zip(reference[6:], slower[6:]))


def test_default_speed_is_the_vacuum_speed():
    default = build_pml_coeffs(**_arguments())
    explicit = build_pml_coeffs(**_arguments(), max_velocity=SPEED_OF_LIGHT)
    tensor = build_pml_coeffs(**_arguments(), max_velocity=torch.tensor(SPEED_OF_LIGHT, dtype=torch.float64))
    for a, b, c in zip(default, explicit, tensor):
        assert torch.equal(a, b) and torch.equal(a, c)


def test_compute_names_share_the_keyword_only_speed():
    for function in (dg.compute, dg.solver.compute, dg.solver.modeling.compute):
        assert function is dg.compute
        parameter = inspect.signature(function).parameters["pml_max_velocity"]
        assert parameter.kind == inspect.Parameter.KEYWORD_ONLY and parameter.default == SPEED_OF_LIGHT


def test_older_builder_name_takes_the_older_material_keywords():
    args = _arguments()
    expected = build_pml_coeffs(**args, max_velocity=1.7e8)
    args["er"] = args.pop("eps_r")
    args["mr"] = args.pop("mu_r")
    for wrapper in (dg.buildpmlcoeffs, dg.solver.buildpmlcoeffs):
        for actual, reference in zip(wrapper(**args, max_velocity=1.7e8), expected):
            assert torch.equal(actual, reference)
    with pytest.raises(TypeError):
        dg.buildpmlcoeffs(**args, eps_r=args["er"])
