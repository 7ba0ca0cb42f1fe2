"""Solves on other streams and from concurrent threads give the bits of one
solve on the default stream (restates the reference's
tests/test_cuda_resource_cache.py:30-131; the reference allows 2e-5 for its
atomic 3D adjoint, this adjoint is a deterministic gather and every result is
compared bit for bit). clear_cuda_resource_cache() between and during calls
changes nothing."""
import threading
import warnings
from concurrent.futures import ThreadPoolExecutor

import pytest
import torch

pytestmark = pytest.mark.gpu


def _run_case(shape=(17, 19), dtype=torch.float32, rhs="stored", offload=True, material=True, device="cuda:0",
              scale=1.0, interval=1, nt=24):
    import deepgpr_amd as dg
    device = torch.device(device)
    three_d = len(shape) == 3
    eps = (torch.arange(int(torch.tensor(shape).prod()), device=device).reshape(shape).float() % 13 * 0.02
           + 4).requires_grad_(material)
    sigma = torch.full(shape, 2e-4, device=device, requires_grad=material)
    dt = 1.5e-11
    source = (scale * dg.ricker(8e8, nt, dt, 1e-10).to(device).reshape(1, nt, 1)).requires_grad_()
    z = 3 if three_d else 0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        out = dg.compute(device=device, dx=0.02, dt=dt, eps_r=eps, sigma=sigma, source_amplitudes=source,
                         source_location=torch.tensor([[[4, 4, z]]], device=device),
                         receiver_location=torch.tensor([[[5, 5, z]]], device=device), pmlthick=2,
                         mode=3 if three_d else 2, fdtd_order=4, source_direction=0 if three_d else 2,
                         use_async_offload=offload, wavefield_storage_dtype=dtype, wavefield_rhs_history=rhs,
                         model_gradient_sampling_interval=interval)
    node = out[-1].grad_fn
    values = [out[-1], *out[1], *out[2], *out[3]]
    values += [v for v in (getattr(node, n, None) for n in ("E_saved", "R_saved", "E_final"))
               if isinstance(v, torch.Tensor)]
    values = [v.detach().cpu().clone() for v in values]
    out[-1].square().sum().backward()
    gradients = [source.grad] + ([eps.grad, sigma.grad] if material else [])
    return values, [g.detach().cpu().clone() for g in gradients]


def _same(expected, actual):
    for left, right in zip(expected, actual):
        assert len(left) == len(right)
        for a, b in zip(left, right):
            assert torch.equal(a, b)


def test_reuse_across_shapes_formats_and_history_layouts(cuda_device):
    import deepgpr_amd as dg
    print("[gpu streams threads] reuse across shapes", flush=True)
    cases = [dict(shape=(17, 19), dtype=torch.float32, rhs="stored"),
             dict(shape=(23, 27), dtype=torch.float16, rhs="stored", interval=2),
             dict(shape=(11, 13, 9), dtype=torch.bfloat16, rhs="stored"),
             dict(shape=(13, 15), dtype=torch.float32, rhs="reconstructed"),
             dict(shape=(19, 21), dtype=torch.float16, rhs="reconstructed"),
             dict(shape=(17, 19), material=False)]
    references = [_run_case(offload=False, **case) for case in cases]
    for _ in range(2):
        for case, expected in zip(cases, references):
            _same(expected, _run_case(**case))
    dg.clear_cuda_resource_cache()


def test_clear_and_a_non_default_stream(cuda_device):
    import deepgpr_amd as dg
    print("[gpu streams threads] non-default stream", flush=True)
    expected = _run_case()
    for _ in range(3):
        dg.clear_cuda_resource_cache()
        with torch.cuda.stream(torch.cuda.Stream()):
            actual = _run_case()
        torch.cuda.synchronize()
        _same(expected, actual)


def test_concurrent_calls_and_clear(cuda_device):
    import deepgpr_amd as dg
    print("[gpu streams threads] two threads on their own streams", flush=True)
    scales = (0.7, 1.3)
    expected = [_run_case(scale=s) for s in scales]
    barrier = threading.Barrier(3)

    def worker(scale):
        with torch.cuda.device(0), torch.cuda.stream(torch.cuda.Stream()):
            barrier.wait(timeout=30)
            result = _run_case(scale=scale)
            torch.cuda.current_stream().synchronize()
            return result

    with ThreadPoolExecutor(max_workers=2) as pool:
        for _ in range(3):
            futures = [pool.submit(worker, s) for s in scales]
            barrier.wait(timeout=30)
            dg.clear_cuda_resource_cache()
            for reference, future in zip(expected, futures):
                _same(reference, future.result(timeout=60))
