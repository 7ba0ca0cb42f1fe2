"""The CPML face profiles are computed once per (thickness, dt, spacing,
reference speed, device, dtype) and copied for every later call."""
import torch

from deepgpr_amd.solver.pml import build_pml_coeffs


def _coeffs(dt, pml=(6, 4, 5, 0, 0, 0), dx=0.02):
    eps = torch.full((10, 12, 1), 4.0)
    return build_pml_coeffs(eps, torch.ones_like(eps), dt, dx, 20, 21, 1, torch.tensor(pml), "cpu", torch.float32)


def test_cached_profiles_are_equal_independent_copies():
    a, b = _coeffs(3e-11), _coeffs(3e-11)
    for x, y in zip(a[6:], b[6:]):
        assert torch.equal(x, y)
        if x.numel():
            assert x.data_ptr() != y.data_ptr()
    a[6].zero_()  # a caller may write into what it got
    assert torch.equal(b[6], _coeffs(3e-11)[6]) and b[6].abs().sum() > 0


def test_profiles_follow_their_arguments():
    a, b = _coeffs(3e-11), _coeffs(2e-11)
    assert not torch.equal(a[6], b[6])
    c = _coeffs(3e-11, dx=0.01)
    assert not torch.equal(a[6], c[6])
    d = _coeffs(3e-11, pml=(6, 4, 5, 0, 0, 0))
    assert torch.equal(a[8], d[8])  # the x-high face (thickness 4)
    assert a[6].shape == (4, 1, 6) and a[8].shape == (4, 1, 4) and a[10].shape == (4, 1, 5)
