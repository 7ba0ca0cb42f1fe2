"""Grid spacing through compute() on the HIP device (restates the reference's
tests/test_numerics.py:499-697): a scalar, a list and a tensor of equal
spacings give the same bits; an anisotropic solve and its adjoint are finite;
one curl update and one source injection scale as 1/d along the axis they
use; the permittivity gradient at a cell matches a central difference with
anisotropic spacing."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _2d_case(dev, nt):
    import deepgpr_amd as dg
    nx, ny = 14, 18
    source = dg.wavelet.ricker(3.0e8, nt, 3.0e-11, 3.0e-9).reshape(1, nt, 1).to(dev)
    return nx, ny, dict(device=dev, dt=3.0e-11, source_amplitudes=source,
                        source_location=torch.tensor([[[6, 7, 0]]], dtype=torch.int32, device=dev),
                        receiver_location=torch.tensor([[[6, 11, 0]]], dtype=torch.int32, device=dev))


def test_scalar_and_equal_axis_spacings_give_the_same_bits(cuda_device):
    import deepgpr_amd as dg
    nx, ny, common = _2d_case(cuda_device, 100)
    args = dict(common, er=torch.full((nx, ny), 4.0, device=cuda_device),
                se=torch.zeros((nx, ny), device=cuda_device), pmlthick=3, fdtd_order=2, mode=2)
    scalar = dg.compute(dx=0.02, **args)[-1]
    vector = dg.compute(dx=[0.02, 0.02, 0.02], **args)[-1]
    tensor = dg.compute(dx=torch.tensor([0.02, 0.02, 0.02], dtype=torch.float64), **args)[-1]
    assert torch.equal(scalar, vector) and torch.equal(scalar, tensor)


def test_anisotropic_forward_and_backward_are_finite(cuda_device):
    import deepgpr_amd as dg
    nx, ny, common = _2d_case(cuda_device, 120)
    er = torch.full((nx, ny), 4.0, device=cuda_device, requires_grad=True)
    receiver = dg.compute(dx=[0.02, 0.015, 0.01], er=er, se=torch.zeros((nx, ny), device=cuda_device), pmlthick=3,
                          fdtd_order=2, mode=2, **common)[-1]
    receiver.square().sum().backward()
    assert bool(torch.isfinite(receiver).all()) and bool(torch.isfinite(er.grad).all())
    assert int(torch.count_nonzero(er.grad[4:-4, 4:-4])) > 0


def test_curl_updates_use_the_spacing_of_their_axis(cuda_device):
    import deepgpr_amd as dg
    dev = cuda_device

    def updated_ez(dy):
        nx, ny = 8, 10
        shape = (1, nx + 1, ny + 1, 2)
        hx = torch.zeros(shape, device=dev)
        hx[:] = torch.arange(ny + 1, dtype=torch.float32, device=dev).reshape(1, 1, ny + 1, 1)
        location = torch.tensor([[[4, 5, 0]]], dtype=torch.int32, device=dev)
        result = dg.compute(device=dev, dx=[0.02, dy, 0.02], dt=3.0e-11, source_amplitudes=torch.zeros((1, 1, 1), device=dev),
                            source_location=location, receiver_location=location,
                            er=torch.full((nx, ny), 4.0, device=dev), se=torch.zeros((nx, ny), device=dev),
                            E=tuple(torch.zeros(shape, device=dev) for _ in range(3)),
                            H=(hx, torch.zeros(shape, device=dev), torch.zeros(shape, device=dev)), pmlthick=0,
                            fdtd_order=2, mode=2)
        return result[1][2][0, 4, 5, 0]

    def updated_ex(dz):
        nx, ny, nz = 6, 7, 8
        shape = (1, nx + 1, ny + 1, nz + 1)
        hy = torch.zeros(shape, device=dev)
        hy[:] = torch.arange(nz + 1, dtype=torch.float32, device=dev).reshape(1, 1, 1, nz + 1)
        location = torch.tensor([[[3, 3, 4]]], dtype=torch.int32, device=dev)
        result = dg.compute(device=dev, dx=[0.02, 0.02, dz], dt=2.0e-11, source_amplitudes=torch.zeros((1, 1, 1), device=dev),
                            source_location=location, receiver_location=location,
                            er=torch.full((nx, ny, nz), 4.0, device=dev), se=torch.zeros((nx, ny, nz), device=dev),
                            E=tuple(torch.zeros(shape, device=dev) for _ in range(3)),
                            H=(torch.zeros(shape, device=dev), hy, torch.zeros(shape, device=dev)), pmlthick=0,
                            source_direction=0, reciever_direction=0, fdtd_order=2, mode=3)
        return result[1][0][0, 3, 3, 4]

    torch.testing.assert_close(updated_ez(0.01), 2.0 * updated_ez(0.02), rtol=1.0e-6, atol=0.0)
    torch.testing.assert_close(updated_ex(0.01), 2.0 * updated_ex(0.02), rtol=1.0e-6, atol=0.0)


def test_source_scaling_uses_the_cell_dimensions(cuda_device):
    import deepgpr_amd as dg
    dev = cuda_device

    def source_sample(dy):
        location = torch.tensor([[[4, 5, 0]]], dtype=torch.int32, device=dev)
        return dg.compute(device=dev, dx=[0.02, dy, 0.01], dt=2.0e-11, source_amplitudes=torch.ones((1, 1, 1), device=dev),
                          source_location=location, receiver_location=location,
                          er=torch.full((8, 10), 4.0, device=dev), se=torch.zeros((8, 10), device=dev), pmlthick=0,
                          source_direction=2, reciever_direction=2, fdtd_order=2, mode=2)[-1][0, 0, 0]

    torch.testing.assert_close(source_sample(0.01), 2.0 * source_sample(0.02), rtol=1.0e-6, atol=0.0)


def test_anisotropic_material_gradient_matches_a_central_difference(cuda_device):
    import deepgpr_amd as dg
    nx, ny, common = _2d_case(cuda_device, 160)
    se = torch.zeros((nx, ny), device=cuda_device)

    def objective(model):
        receiver = dg.compute(dx=[0.02, 0.015, 0.01], er=model, se=se, pmlthick=3, model_gradient_sampling_interval=1,
                              wavefield_storage_dtype=torch.float32, fdtd_order=2, mode=2, **common)[-1]
        return 0.5 * receiver.square().sum()

    model = torch.full((nx, ny), 4.0, device=cuda_device, requires_grad=True)
    objective(model).backward()
    direction = torch.zeros_like(model)
    direction[6, 9] = 1.0
    with torch.no_grad():
        fd = (objective(model.detach() + 1.0e-2 * direction) - objective(model.detach() - 1.0e-2 * direction)) / 2.0e-2
    adjoint = model.grad[6, 9]
    assert abs(float(adjoint - fd)) / max(abs(float(adjoint)), abs(float(fd)), 1.0e-20) < 0.02
