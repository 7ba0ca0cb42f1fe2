"""Dot-product (adjoint) tests of the whole state map through compute() on
the HIP device, and run-to-run reproducibility of the 2D Ez-TM adjoint
(restates the reference's tests/test_discrete_adjoint.py:65-195, 362-432):
for random initial states x and random output duals y, <F x, y> must equal
<x, F^T y> to 2e-4 in float32 for one step with every CPML face, orders
2/4/8, symmetric and asymmetric faces, on the general 3D kernels and the 2D
Ez-TM kernels, whose inactive components must get zero cotangents."""
import pytest
import torch

pytestmark = pytest.mark.gpu

_TM2D_INACTIVE = (0, 1, 5) + tuple(6 + index for index in (0, 3, 4, 7, 8, 11, 12, 15))


def _state_problem(dev, order, pml):
    import deepgpr_amd as dg
    torch.manual_seed(1100 + order + (pml if isinstance(pml, int) else sum(pml)))
    shape = (9, 10, 8)
    nx, ny, nz = shape
    x = torch.linspace(0.0, 1.0, nx, device=dev)[:, None, None]
    y = torch.linspace(0.0, 1.0, ny, device=dev)[None, :, None]
    z = torch.linspace(0.0, 1.0, nz, device=dev)[None, None, :]
    eps_r = 3.5 + 0.7 * x + 0.2 * y + 0.1 * z
    sigma = 1.0e-4 + 1.0e-4 * (x + y + z) / 3.0
    mu_r = 1.0 + 0.15 * x + 0.0 * y + 0.05 * z
    location = torch.tensor([[[max(2, s // 2) for s in shape]]], dtype=torch.int32, device=dev)
    source = torch.zeros((1, 1, 1), dtype=torch.float32, device=dev)
    spacing, dt = (0.020, 0.017, 0.013), 1.0e-11
    e0, h0, p0 = dg.checkpoint_initial_field(device=dev, dx=spacing, dt=dt, source_amplitudes=source,
                                             source_location=location, receiver_location=location, er=eps_r,
                                             se=sigma, mr=mu_r, pmlthick=pml, fdtd_order=order)
    base = [torch.randn_like(t) * 1.0e-3 for t in (*e0, *h0, *p0)]
    state = [t.clone().requires_grad_(True) for t in base]
    result = dg.compute(device=dev, dx=spacing, dt=dt, source_amplitudes=source, source_location=location,
                        receiver_location=location, er=eps_r, se=sigma, mr=mu_r, E=tuple(state[:3]),
                        H=tuple(state[3:6]), PML=tuple(state[6:]), pmlthick=pml, source_direction=2,
                        reciever_direction=2, fdtd_order=order, mode=3, model_gradient_sampling_interval=1,
                        wavefield_storage_dtype=torch.float32, debug=True)
    outputs = [*result[1], *result[2], *result[3], result[-1]]
    duals = [torch.randn_like(t) for t in outputs]
    output_dot = sum((v * d).sum() for v, d in zip(outputs, duals))
    output_dot.backward()
    input_dot = sum((b * v.grad).sum() for b, v in zip(base, state))
    out, inp = float(output_dot.detach()), float(input_dot.detach())
    return abs(out - inp) / max(abs(out), abs(inp), 1.0e-20)


def _state_problem_2d_tm(dev, order, pml):
    import deepgpr_amd as dg
    torch.manual_seed(3100 + order + (pml if isinstance(pml, int) else sum(pml)))
    nx, ny = 11, 12
    x = torch.linspace(0.0, 1.0, nx, device=dev)[:, None]
    y = torch.linspace(0.0, 1.0, ny, device=dev)[None, :]
    eps_r = 3.5 + 0.7 * x + 0.2 * y
    sigma = 1.0e-4 + 1.0e-4 * (x + y) / 2.0
    mu_r = 1.0 + 0.15 * x + 0.05 * y
    location = torch.tensor([[[5, 6, 0]]], dtype=torch.int32, device=dev)
    source = torch.zeros((1, 1, 1), dtype=torch.float32, device=dev)
    spacing, dt = (0.020, 0.017, 0.013), 1.0e-11
    e0, h0, p0 = dg.checkpoint_initial_field(device=dev, dx=spacing, dt=dt, source_amplitudes=source,
                                             source_location=location, receiver_location=location, er=eps_r,
                                             se=sigma, mr=mu_r, pmlthick=pml, fdtd_order=order)
    base = [torch.randn_like(t) * 1.0e-3 for t in (*e0, *h0, *p0)]
    for index in _TM2D_INACTIVE:
        base[index] = torch.zeros_like(base[index])
    state = [t.clone().requires_grad_(True) for t in base]
    result = dg.compute(device=dev, dx=spacing, dt=dt, source_amplitudes=source, source_location=location,
                        receiver_location=location, er=eps_r, se=sigma, mr=mu_r, E=tuple(state[:3]),
                        H=tuple(state[3:6]), PML=tuple(state[6:]), pmlthick=pml, source_direction=2,
                        reciever_direction=2, fdtd_order=order, mode=2)
    tm2d = bool(result[-1].grad_fn.config.tm2d_fast_path)
    outputs = [*result[1], *result[2], *result[3], result[-1]]
    duals = [torch.randn_like(t) for t in outputs]
    output_dot = sum((v * d).sum() for v, d in zip(outputs, duals))
    output_dot.backward()
    input_dot = sum((b * v.grad).sum() for k, (b, v) in enumerate(zip(base, state)) if k not in _TM2D_INACTIVE)
    inactive = max(float(state[k].grad.abs().max()) if state[k].numel() else 0.0 for k in _TM2D_INACTIVE)
    out, inp = float(output_dot.detach()), float(input_dot.detach())
    return abs(out - inp) / max(abs(out), abs(inp), 1.0e-20), tm2d, inactive


@pytest.mark.parametrize("order", [2, 4, 8])
def test_cpml_single_step_dot_products(cuda_device, order):
    print(f"[gpu dot products] order {order}, CPML 2", flush=True)
    assert _state_problem(cuda_device, order, 2) < 2.0e-4


@pytest.mark.parametrize("order", [2, 4, 8])
def test_asymmetric_cpml_dot_products(cuda_device, order):
    print(f"[gpu dot products] order {order}, asymmetric CPML", flush=True)
    for pml in ([1, 2, 3, 1, 2, 3], [3, 0, 1, 2, 0, 2]):
        assert _state_problem(cuda_device, order, pml) < 2.0e-4, pml
    error, tm2d, inactive = _state_problem_2d_tm(cuda_device, order, [3, 1, 0, 2])
    assert tm2d and error < 2.0e-4 and inactive == 0.0


def test_2d_tm_adjoint_is_bitwise_reproducible(cuda_device):
    import deepgpr_amd as dg
    print("[gpu dot products] 2D Ez-TM adjoint reproducibility", flush=True)
    dev = cuda_device

    def run(shape, mode, sources, receivers, pml):
        torch.manual_seed(5)
        nt = 50
        eps_r = (4.0 + torch.rand(shape, device=dev)).requires_grad_(True)
        sigma = (1.0e-4 + 1.0e-4 * torch.rand(shape, device=dev)).requires_grad_(True)
        source = torch.randn((sources.shape[1], nt, 1), device=dev).requires_grad_(True)
        e0, h0, p0 = dg.checkpoint_initial_field(device=dev, dx=0.02, dt=1.0e-11, source_amplitudes=source.detach(),
                                                 source_location=sources, receiver_location=receivers,
                                                 er=eps_r.detach(), se=sigma.detach(), pmlthick=pml, fdtd_order=4)
        states = [t.clone().requires_grad_(True) for t in (*e0, *h0, *p0)]
        result = dg.compute(device=dev, dx=0.02, dt=1.0e-11, source_amplitudes=source, source_location=sources,
                            receiver_location=receivers, eps_r=eps_r, sigma=sigma, pmlthick=pml, fdtd_order=4,
                            mode=mode, E=tuple(states[:3]), H=tuple(states[3:6]), PML=tuple(states[6:]))
        assert result[-1].grad_fn.config.tm2d_fast_path
        loss = result[-1].square().sum() + sum(t.square().sum() for t in (*result[1], *result[2]))
        loss.backward()
        return [eps_r.grad, sigma.grad, source.grad] + [t.grad for t in states]

    coincident = torch.tensor([[[6, 5, 0], [6, 5, 0]], [[8, 7, 0], [8, 7, 0]]], dtype=torch.int32, device=dev)
    receivers = torch.tensor([[[6, 9, 0], [6, 9, 0], [7, 11, 0]], [[8, 10, 0], [8, 10, 0], [5, 12, 0]]],
                             dtype=torch.int32, device=dev)
    for mode, pml in ((2, [2, 3, 1, 2]), (3, [3, 0, 2, 1])):
        first = run((14, 16), mode, coincident, receivers, pml)
        second = run((14, 16), mode, coincident, receivers, pml)
        for a, b in zip(first, second):
            assert torch.equal(a, b), mode
