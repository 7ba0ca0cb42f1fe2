"""Adjoint cases of the reference's tests/test_discrete_adjoint.py:204-284,
317-335, 475-547 and 585-800 through compute() on the HIP device: coincident
sources and receivers, Taylor checks of full-model material directions
without and with CPML, the choice of the 2D Ez-TM kernels and their identity
with the general kernels, finite results for a large all-face CPML with FP16
histories and for an order-8 asymmetric CPML with an incomplete sampling
block, the older keyword names, the source-waveform gradient, the weight
of an incomplete sampling block, and a long 3D CPML backward
(tests/test_numerics.py:717-754)."""
import contextlib

import pytest
import torch

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def _general_kernels():
    from deepgpr_amd.solver import modeling
    original = modeling._tm2d_applies
    modeling._tm2d_applies = lambda *args, **kwargs: False
    try:
        yield
    finally:
        modeling._tm2d_applies = original


def test_coincident_sources_and_receivers_accumulate(cuda_device):
    import deepgpr_amd as dg
    print("[gpu adjoint cases] coincident sources and receivers", flush=True)
    dev, nt = cuda_device, 48
    eps_r = torch.full((12, 14), 4.0, device=dev)
    sigma = torch.zeros_like(eps_r)
    first = torch.linspace(-0.5, 0.8, nt, device=dev)
    second = torch.linspace(0.2, -0.4, nt, device=dev)
    one_source = torch.tensor([[[5, 5, 0]]], dtype=torch.int32, device=dev)
    one_receiver = torch.tensor([[[5, 8, 0]]], dtype=torch.int32, device=dev)
    common = dict(device=dev, dx=0.02, dt=3.0e-11, receiver_location=one_receiver, er=eps_r, se=sigma, pmlthick=0,
                  fdtd_order=2, mode=2)
    pair = dg.compute(source_amplitudes=torch.stack((first, second), dim=0).unsqueeze(-1),
                      source_location=one_source.repeat(1, 2, 1), **common)[-1]
    summed = dg.compute(source_amplitudes=(first + second).reshape(1, nt, 1), source_location=one_source, **common)[-1]
    torch.testing.assert_close(pair, summed, rtol=2.0e-5, atol=1.0e-6)

    def receiver_gradient(receivers):
        model = eps_r.detach().clone().requires_grad_(True)
        data = dg.compute(device=dev, dx=0.02, dt=3.0e-11, source_amplitudes=(first + second).reshape(1, nt, 1),
                          source_location=one_source, receiver_location=receivers, er=model, se=sigma, pmlthick=0,
                          fdtd_order=2, mode=2, model_gradient_sampling_interval=1,
                          wavefield_storage_dtype=torch.float32)[-1]
        data.sum().backward()
        return model.grad

    single = receiver_gradient(one_receiver)
    duplicate = receiver_gradient(one_receiver.repeat(1, 2, 1))
    torch.testing.assert_close(duplicate, 2.0 * single, rtol=2.0e-4, atol=1.0e-6)


def _material_case(dev, pml):
    import deepgpr_amd as dg
    torch.manual_seed(2100 + pml)
    nx, ny, nt = 14, 17, 100
    x = torch.linspace(0.0, 1.0, nx)[:, None]
    y = torch.linspace(0.0, 1.0, ny)[None, :]
    eps_base = (4.0 + 0.4 * x + 0.2 * y).to(dev)
    sigma_base = (2.5e-4 + 0.5e-4 * x + 0.3e-4 * y).to(dev)
    mu_r = (1.0 + 0.1 * x + 0.05 * y).to(dev)
    source = dg.wavelet.ricker(3.5e8, nt, 2.5e-11, 2.0e-9).reshape(1, nt, 1).to(dev)
    arguments = dict(device=dev, dx=(0.020, 0.016, 0.012), dt=2.5e-11, source_amplitudes=source,
                     source_location=torch.tensor([[[6, 6, 0]]], dtype=torch.int32, device=dev),
                     receiver_location=torch.tensor([[[6, 10, 0], [7, 12, 0]]], dtype=torch.int32, device=dev),
                     mr=mu_r, pmlthick=pml, fdtd_order=4, mode=2, model_gradient_sampling_interval=1,
                     wavefield_storage_dtype=torch.float32)
    with torch.no_grad():
        reference = dg.compute(er=eps_base, se=sigma_base, **arguments)[-1]
    scale = reference.double().abs().max().clamp_min(1.0e-12)
    target = 0.85 * reference.double()

    def objective(eps_r, sigma):
        return 0.5 * ((dg.compute(er=eps_r, se=sigma, **arguments)[-1].double() - target) / scale).square().sum()

    eps_r = eps_base.clone().requires_grad_(True)
    sigma = sigma_base.clone().requires_grad_(True)
    loss = objective(eps_r, sigma)
    loss.backward()
    # directions from the CPU generator, as the reference draws them on the CPU
    d_eps = torch.randn(eps_base.shape)
    d_eps = (d_eps / d_eps.abs().max().clamp_min(1.0e-12) * 0.15).to(dev)
    d_sig = torch.randn(sigma_base.shape)
    d_sig = (d_sig / d_sig.abs().max().clamp_min(1.0e-12) * 2.0e-4).to(dev)
    rows = {}
    for name, (de, ds) in {"eps_r": (d_eps, torch.zeros_like(d_sig)), "sigma": (torch.zeros_like(d_eps), d_sig),
                           "joint": (d_eps, d_sig)}.items():
        adjoint = float((eps_r.grad.double() * de.double()).sum() + (sigma.grad.double() * ds.double()).sum())
        case = []
        for step in (1.0, 0.5, 0.25, 0.125, 0.0625, 0.03125, 0.015625):
            with torch.no_grad():
                plus = float(objective(eps_base + step * de, sigma_base + step * ds))
                minus = float(objective(eps_base - step * de, sigma_base - step * ds))
            fd = (plus - minus) / (2.0 * step)
            error = abs(adjoint - fd) / max(abs(adjoint), abs(fd), 1.0e-20)
            case.append((step, error, abs(plus - float(loss.detach()) - step * adjoint)))
        rows[name] = case
    return rows


@pytest.mark.parametrize("pml", [0, 3])
def test_material_taylor_checks(cuda_device, pml):
    print(f"[gpu adjoint cases] Taylor checks, CPML {pml}", flush=True)
    for name, case in _material_case(cuda_device, pml).items():
        errors = [r[1] for r in case]
        remainders = [r[2] for r in case]
        assert min(errors) < 1.5e-2, (name, errors)
        if name == "sigma":
            assert min(remainders) < 2.0e-6, (name, remainders)
        else:
            assert remainders[-1] < remainders[0] * 0.2, (name, remainders)


def test_2d_tm_kernel_choice_and_identity_with_the_general_kernels(cuda_device):
    import deepgpr_amd as dg
    print("[gpu adjoint cases] 2D Ez-TM kernel choice", flush=True)
    dev = cuda_device
    nx, ny, nt = 16, 18, 40
    eps_r = torch.full((nx, ny), 4.0, device=dev)
    sigma = torch.full_like(eps_r, 2.0e-4)
    source = dg.wavelet.ricker(4.0e8, nt, 2.5e-11, 1.5e-9).reshape(1, nt, 1).to(dev)
    location = torch.tensor([[[7, 6, 0]]], dtype=torch.int32, device=dev)
    receivers = torch.tensor([[[7, 12, 0], [9, 13, 0]]], dtype=torch.int32, device=dev)
    common = dict(device=dev, dx=0.02, dt=2.5e-11, source_amplitudes=source, source_location=location,
                  receiver_location=receivers, eps_r=eps_r, sigma=sigma, pmlthick=[3, 2, 2, 4], fdtd_order=4, mode=2)
    fast = dg.compute(**common)
    with _general_kernels():
        general = dg.compute(**common)
    assert torch.equal(fast[-1], general[-1])
    for a, b in zip((*fast[1], *fast[2]), (*general[1], *general[2])):
        assert torch.equal(a, b)

    def selected(**extra):
        eps = eps_r.clone().requires_grad_(True)
        return dg.compute(**{**common, "eps_r": eps, **extra})[-1].grad_fn.config.tm2d_fast_path

    assert selected()
    assert not selected(source_direction=0, receiver_component=0, mode=3)
    e0, h0, p0 = dg.checkpoint_initial_field(device=dev, dx=0.02, dt=2.5e-11, source_amplitudes=source,
                                             source_location=location, receiver_location=receivers, er=eps_r,
                                             se=sigma, pmlthick=[3, 2, 2, 4], fdtd_order=4)
    assert selected(E=e0, H=h0, PML=p0)
    nonzero_ey = (e0[0].clone(), e0[1].clone(), e0[2].clone())
    nonzero_ey[1][0, 5, 5, 0] = 1.0e-3
    assert not selected(E=nonzero_ey)
    nonzero_phi = [t.clone() for t in p0]
    nonzero_phi[0].fill_(1.0e-6)
    assert not selected(PML=nonzero_phi)


def test_large_all_face_cpml_fp16_history(cuda_device):
    import deepgpr_amd as dg
    print("[gpu adjoint cases] 120x120x100, CPML 10, FP16, S 3", flush=True)
    nx, ny, nz, nt = 120, 120, 100, 5
    eps_r = torch.full((nx, ny, nz), 4.0, dtype=torch.float32, device=cuda_device, requires_grad=True)
    sigma = torch.full_like(eps_r, 1.0e-4)
    source = torch.zeros((1, nt, 1), dtype=torch.float32, device=cuda_device)
    source[0, 0, 0] = 1.0
    result = dg.compute(device=cuda_device, dx=(0.020, 0.017, 0.014), dt=1.0e-11, source_amplitudes=source,
                        source_location=torch.tensor([[[60, 60, 50]]], dtype=torch.int32),
                        receiver_location=torch.tensor([[[60, 62, 50]]], dtype=torch.int32), er=eps_r, se=sigma,
                        pmlthick=10, fdtd_order=2, mode=3, model_gradient_sampling_interval=3,
                        wavefield_storage_dtype=torch.float16, use_async_offload=True)
    loss = result[-1].square().sum() + sum(f.square().mean() for f in (*result[1], *result[2]))
    loss.backward()
    torch.cuda.synchronize()
    assert bool(torch.isfinite(result[-1]).all())
    assert eps_r.grad is not None and bool(torch.isfinite(eps_r.grad).all())


def test_order_8_asymmetric_cpml_and_incomplete_sampling_block(cuda_device):
    import deepgpr_amd as dg
    print("[gpu adjoint cases] order 8, asymmetric CPML, BF16, S 4 over 11 steps", flush=True)
    shape, nt = (31, 29, 27), 11
    eps_r = torch.full(shape, 4.0, device=cuda_device, requires_grad=True)
    sigma = torch.full_like(eps_r, 2.0e-4, requires_grad=True)
    source = torch.linspace(-0.3, 0.7, nt, device=cuda_device).reshape(1, nt, 1)
    receiver = dg.compute(device=cuda_device, dx=(0.020, 0.016, 0.013), dt=1.0e-11, source_amplitudes=source,
                          source_location=torch.tensor([[[15, 14, 13]]], dtype=torch.int32, device=cuda_device),
                          receiver_location=torch.tensor([[[16, 15, 14]]], dtype=torch.int32, device=cuda_device),
                          er=eps_r, se=sigma, pmlthick=[3, 5, 4, 2, 6, 3], fdtd_order=8, mode=3,
                          model_gradient_sampling_interval=4, wavefield_storage_dtype=torch.bfloat16,
                          use_async_offload=True)[-1]
    receiver.square().sum().backward()
    torch.cuda.synchronize()
    assert bool(torch.isfinite(receiver).all())
    assert bool(torch.isfinite(eps_r.grad).all()) and bool(torch.isfinite(sigma.grad).all())


def test_older_and_current_keyword_names_give_the_same_bits(cuda_device):
    # tests/test_discrete_adjoint.py:317-335
    import deepgpr_amd as dg
    eps_r = torch.full((10, 12), 4.0, device=cuda_device)
    sigma = torch.zeros_like(eps_r)
    location = torch.tensor([[[5, 6, 0]]], dtype=torch.int32, device=cuda_device)
    common = dict(device=cuda_device, dx=0.02, dt=3.0e-11,
                  source_amplitudes=torch.linspace(-0.5, 0.5, 20, device=cuda_device).reshape(1, 20, 1),
                  source_location=location, receiver_location=location, pmlthick=0, fdtd_order=2, mode=2)
    old = dg.compute(er=eps_r, se=sigma, reciever_direction=2, **common)[-1]
    new = dg.compute(eps_r=eps_r, sigma=sigma, receiver_component=2, **common)[-1]
    assert torch.equal(old, new)


def test_source_waveform_gradient_matches_finite_differences(cuda_device):
    # tests/test_discrete_adjoint.py:475-508
    import math
    import deepgpr_amd as dg
    print("[gpu adjoint cases] source waveform gradient", flush=True)
    nx, ny, nt = 12, 14, 60
    arguments = dict(device=cuda_device, dx=0.02, dt=3.0e-11,
                     source_location=torch.tensor([[[5, 5, 0]]], dtype=torch.int32, device=cuda_device),
                     receiver_location=torch.tensor([[[5, 8, 0]]], dtype=torch.int32, device=cuda_device),
                     er=torch.full((nx, ny), 4.0, device=cuda_device),
                     se=torch.zeros((nx, ny), device=cuda_device), pmlthick=0, fdtd_order=2, mode=2)
    base = dg.wavelet.ricker(4.0e8, nt, 3.0e-11, 2.0e-9).reshape(1, nt, 1).to(cuda_device)

    def objective(source):
        return 0.5 * dg.compute(source_amplitudes=source, **arguments)[-1].square().sum()

    source = base.clone().requires_grad_(True)
    objective(source).backward()
    direction = torch.randn(source.shape, generator=torch.Generator().manual_seed(41)).to(cuda_device)
    direction /= direction.norm()
    adjoint = float((source.grad * direction).sum())
    best = math.inf
    for step in (2.0e-3, 1.0e-3, 5.0e-4):
        with torch.no_grad():
            fd = float((objective(base + step * direction) - objective(base - step * direction)) / (2.0 * step))
        best = min(best, abs(adjoint - fd) / max(abs(adjoint), abs(fd), 1.0e-20))
    assert best < 8.0e-3


def test_incomplete_sampling_block_uses_its_actual_length(cuda_device):
    # tests/test_discrete_adjoint.py:523-547: an interval of nt and of nt + 5
    # both save only step 0, whose weight is the block's actual length
    import deepgpr_amd as dg
    nx, ny, nt = 12, 14, 11
    eps_base = torch.full((nx, ny), 4.0, device=cuda_device)
    sigma = torch.full_like(eps_base, 2.0e-4)
    source = torch.linspace(-0.5, 0.8, nt, device=cuda_device).reshape(1, nt, 1)

    def gradient(interval):
        eps_r = eps_base.clone().requires_grad_(True)
        receiver = dg.compute(device=cuda_device, dx=0.02, dt=3.0e-11, source_amplitudes=source,
                              source_location=torch.tensor([[[5, 5, 0]]], dtype=torch.int32, device=cuda_device),
                              receiver_location=torch.tensor([[[5, 7, 0]]], dtype=torch.int32, device=cuda_device),
                              er=eps_r, se=sigma, pmlthick=0, fdtd_order=2, mode=2,
                              model_gradient_sampling_interval=interval, wavefield_storage_dtype=torch.float32)[-1]
        receiver.square().sum().backward()
        return eps_r.grad

    assert torch.equal(gradient(nt), gradient(nt + 5))


def test_long_3d_cpml_backward_remains_finite(cuda_device):
    # tests/test_numerics.py:717-754: 1500 steps, sampling interval 10
    import deepgpr_amd as dg
    print("[gpu adjoint cases] long 3D CPML backward", flush=True)
    nx, ny, nz, nt = 20, 24, 24, 1500
    er = torch.full((nx, ny, nz), 4.0, dtype=torch.float32, device=cuda_device, requires_grad=True)
    se = torch.zeros_like(er, requires_grad=True)
    source = dg.wavelet.ricker(5.0e8, nt, 2.0e-11, 2.0e-9).reshape(1, nt, 1).to(cuda_device)
    result = dg.compute(device=cuda_device, dx=0.02, dt=2.0e-11, source_amplitudes=source,
                        source_location=torch.tensor([[[10, 7, 12]]], dtype=torch.int32, device=cuda_device),
                        receiver_location=torch.tensor([[[8, 7, 10], [12, 7, 14]]], dtype=torch.int32,
                                                       device=cuda_device),
                        er=er, se=se, pmlthick=5, model_gradient_sampling_interval=10,
                        wavefield_storage_dtype=torch.float32, fdtd_order=2, mode=3, debug=True)
    receiver = result[-1]
    receiver.square().mean().backward()
    assert bool(torch.isfinite(receiver).all())
    assert er.grad is not None and bool(torch.isfinite(er.grad).all()) and er.grad.shape == er.shape
    assert se.grad is not None and bool(torch.isfinite(se.grad).all()) and se.grad.shape == se.shape
    assert int(torch.count_nonzero(er.grad[:6])) > 0
