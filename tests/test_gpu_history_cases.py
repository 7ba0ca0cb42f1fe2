"""History options through compute() on the HIP device (restates the
reference's tests/test_wavefield_compression.py:32-137 and 205-255 and
tests/test_cuda_conversion_backends.py:131-210): no history, the INT8
reduction backends, partial INT8 tiles in 3D, and the FP16/BF16 conversion
backends."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _forward_history_case(dev, save_wavefield_history):
    import deepgpr_amd as dg
    nx, ny, nt = 18, 21, 48
    eps_r = torch.full((nx, ny), 4.0, device=dev, requires_grad=True)
    sigma = torch.full((nx, ny), 2.0e-4, device=dev, requires_grad=True)
    source = dg.wavelet.ricker(3.5e8, nt, 2.5e-11, 2.0e-9, device=dev).reshape(1, nt, 1)
    return dg.compute(device=dev, dx=(0.020, 0.016, 0.012), dt=2.5e-11, source_amplitudes=source,
                      source_location=torch.tensor([[[8, 7, 0]]], dtype=torch.int32, device=dev),
                      receiver_location=torch.tensor([[[7, 12, 0], [9, 14, 0]]], dtype=torch.int32, device=dev),
                      eps_r=eps_r, sigma=sigma, pmlthick=3, fdtd_order=4, mode=2,
                      save_wavefield_history=save_wavefield_history)


def test_forward_only_receivers_equal_those_of_a_history_run(cuda_device):
    print("[gpu history cases] no history", flush=True)
    reference = _forward_history_case(cuda_device, True)
    fdtd_only = _forward_history_case(cuda_device, False)
    assert reference[0].numel() > 0 and fdtd_only[0].numel() == 0
    assert torch.equal(fdtd_only[-1].detach(), reference[-1].detach())  # the reference allows 1e-7


def test_forward_only_backward_fails_explicitly(cuda_device):
    result = _forward_history_case(cuda_device, False)
    with pytest.raises(RuntimeError, match="Forward wavefield history is disabled; adjoint/model-gradient "
                                           "backward is unavailable"):
        result[-1].square().mean().backward()


def _run_2d(dev, compression, backend="current"):
    import deepgpr_amd as dg
    nx, ny, nt = 18, 21, 96
    x = torch.linspace(0.0, 1.0, nx, device=dev)[:, None]
    y = torch.linspace(0.0, 1.0, ny, device=dev)[None, :]
    eps_r = (4.0 + 0.25 * x + 0.15 * y).requires_grad_(True)
    sigma = (2.0e-4 + 0.4e-4 * x + 0.2e-4 * y).requires_grad_(True)
    source = dg.wavelet.ricker(3.5e8, nt, 2.5e-11, 2.0e-9, device=dev).reshape(1, nt, 1)
    kwargs = dict(device=dev, dx=(0.020, 0.016, 0.012), dt=2.5e-11, source_amplitudes=source,
                  source_location=torch.tensor([[[8, 7, 0]]], dtype=torch.int32, device=dev),
                  receiver_location=torch.tensor([[[7, 12, 0], [9, 14, 0]]], dtype=torch.int32, device=dev),
                  eps_r=eps_r, sigma=sigma, pmlthick=3, fdtd_order=4, mode=2, model_gradient_sampling_interval=1,
                  wavefield_storage_dtype=torch.float32, wavefield_compression=compression,
                  wavefield_compression_block_size=(8, 8) if compression == "int8" else None,
                  int8_reduction_backend=backend)
    result = dg.compute(**kwargs)
    data = result[-1]
    scale = data.detach().abs().max().clamp_min(1.0e-12)
    (0.5 * ((data - 0.87 * data.detach()) / scale).square().sum()).backward()
    return {"history": result[0].detach(), "data": data.detach(), "eps": eps_r.grad, "sigma": sigma.grad,
            "kwargs": kwargs}


def test_int8_reduction_backends_are_bit_exact(cuda_device):
    print("[gpu history cases] INT8 reduction backends", flush=True)
    current = _run_2d(cuda_device, "int8", "current")
    for backend in ("cub_block", "warp_shuffle"):
        candidate = _run_2d(cuda_device, "int8", backend)
        # the reference bounds the gradients by 1e-6 (its adjoint adds with
        # atomics); this adjoint is a deterministic gather, so they are equal
        for key in ("history", "data", "eps", "sigma"):
            assert torch.equal(candidate[key], current[key]), (backend, key)


def test_int8_specialised_reductions_need_64_voxels(cuda_device):
    import deepgpr_amd as dg
    kwargs = dict(_run_2d(cuda_device, "int8", "current")["kwargs"])
    kwargs["eps_r"] = kwargs["eps_r"].detach()
    kwargs["sigma"] = kwargs["sigma"].detach()
    kwargs["wavefield_compression_block_size"] = (4, 8)
    kwargs["int8_reduction_backend"] = "cub_block"
    with pytest.raises(ValueError, match="exactly 64 voxels"):
        dg.compute(**kwargs)


def test_3d_partial_int8_tiles(cuda_device):
    import deepgpr_amd as dg
    print("[gpu history cases] 3D partial INT8 tiles", flush=True)
    shape, nt = (9, 10, 7), 8
    eps_r = torch.full(shape, 4.0, device=cuda_device, requires_grad=True)
    sigma = torch.full(shape, 2.0e-4, device=cuda_device, requires_grad=True)
    source = torch.zeros((1, nt, 1), device=cuda_device)
    source[0, 1, 0] = 1.0
    location = torch.tensor([[[4, 5, 3]]], dtype=torch.int32, device=cuda_device)
    result = dg.compute(device=cuda_device, dx=(0.020, 0.018, 0.016), dt=1.0e-11, source_amplitudes=source,
                        source_location=location, receiver_location=location, eps_r=eps_r, sigma=sigma, pmlthick=1,
                        mode=3, wavefield_compression="int8", wavefield_compression_block_size=(4, 4, 4))
    result[-1].square().sum().backward()
    torch.cuda.synchronize(cuda_device)
    assert result[0].dtype == torch.int8
    assert bool(torch.isfinite(eps_r.grad).all()) and bool(eps_r.grad.abs().sum() > 0)


def _run_conversion(dev, storage_dtype, backend):
    # tests/test_cuda_conversion_backends.py:131-174
    import deepgpr_amd as dg
    nx, ny, nt = 24, 28, 120
    x = torch.linspace(0.0, 1.0, nx, device=dev)[:, None]
    y = torch.linspace(0.0, 1.0, ny, device=dev)[None, :]
    eps_r = (4.0 + 0.35 * x + 0.15 * y).requires_grad_(True)
    sigma = (2.0e-4 + 0.3e-4 * x + 0.2e-4 * y).requires_grad_(True)
    source = dg.wavelet.ricker(3.5e8, nt, 2.5e-11, 2.0e-9, device=dev).reshape(1, nt, 1)
    result = dg.compute(device=dev, dx=(0.020, 0.016, 0.012), dt=2.5e-11, source_amplitudes=source,
                        source_location=torch.tensor([[[10, 8, 0]]], dtype=torch.int32, device=dev),
                        receiver_location=torch.tensor([[[8, 14, 0], [12, 18, 0]]], dtype=torch.int32, device=dev),
                        eps_r=eps_r, sigma=sigma, pmlthick=4, fdtd_order=4, mode=2, model_gradient_sampling_interval=1,
                        wavefield_storage_dtype=storage_dtype, wavefield_conversion_backend=backend)
    data = result[-1]
    (0.5 * (data / data.detach().abs().max().clamp_min(1.0e-12)).square().sum()).backward()
    return {"history": result[0].detach(), "data": data.detach(), "eps": eps_r.grad, "sigma": sigma.grad}


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_conversion_backends_give_the_same_solve(cuda_device, dtype):
    # the reference allows 5e-7 between its backends; these share one
    # round-to-nearest-even conversion, so every quantity is equal
    print(f"[gpu history cases] conversion backends {dtype}", flush=True)
    legacy = _run_conversion(cuda_device, dtype, "legacy")
    for backend in ("native_scalar", "native_vec2"):
        candidate = _run_conversion(cuda_device, dtype, backend)
        for name in legacy:
            assert torch.equal(candidate[name], legacy[name]), (backend, name)


def test_conversion_backend_is_refused_where_it_does_not_apply(cuda_device):
    import deepgpr_amd as dg
    with pytest.raises(ValueError, match="only selectable"):
        dg.compute(device=cuda_device, dx=0.02, dt=1.0e-11, source_amplitudes=torch.zeros((1, 2, 1), device=cuda_device),
                   source_location=torch.tensor([[[2, 2, 0]]], dtype=torch.int32, device=cuda_device),
                   receiver_location=torch.tensor([[[2, 2, 0]]], dtype=torch.int32, device=cuda_device),
                   eps_r=torch.full((6, 6), 4.0, device=cuda_device), sigma=torch.full((6, 6), 2.0e-4, device=cuda_device),
                   pmlthick=1, wavefield_storage_dtype=torch.float32, wavefield_conversion_backend="native_scalar")
