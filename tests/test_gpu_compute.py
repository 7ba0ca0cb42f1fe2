"""End-to-end checks of the public compute() API on the GPU.

The native calls made by the autograd layer are captured (inputs before,
outputs after) and replayed through the CPU oracle, so the whole host path -
model extension, CPML profiles, location shift, history allocation and the
storage_type bit field - is checked bit for bit against the oracle."""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.dirname(__file__))

pytestmark = pytest.mark.gpu


from gpu_helpers import capture as _capture  # noqa: E402
from gpu_helpers import np_of as _np  # noqa: E402
from gpu_helpers import problem as _problem  # noqa: E402


def _c1_inputs(dev):
    from deepgpr_amd import synthetic
    w = synthetic.c1_small_2d()
    nt = 400
    return w, dict(dx=w.dx, dt=w.dt, source_amplitudes=w.source_amplitudes[:, :nt].to(dev),
                   source_location=w.source_location.to(dev), receiver_location=w.receiver_location.to(dev))


@pytest.mark.parametrize("acc", ["float32", "float64"])
def test_compute_c1_matches_oracle(oracle_built, cuda_device, monkeypatch, acc):
    import deepgpr_amd as dg
    h = oracle_built
    calls = _capture(monkeypatch)
    w, kw = _c1_inputs(cuda_device)
    eps = w.eps_r.to(cuda_device).requires_grad_(True)
    sig = w.sigma.to(cuda_device).requires_grad_(True)
    out = dg.compute(cuda_device, eps_r=eps, sigma=sig, gradient_accumulator=acc, **kw)
    rx = out[-1]
    (rx * rx).sum().backward()
    (_, fin, fout), (_, bin_, bout) = calls
    assert fin["storage_type"] & (1 << 6), "the 2D Ez case must take the TM fast path"
    p = _problem(h, fin, bin_)
    cpu = h.run_cpu(p, "oracle")
    assert np.array_equal(_np(fout["receiver_data"]), cpu["rx"])
    assert np.array_equal(_np(fout["E_saved"]).view(np.uint8).ravel(), cpu["E_saved"])
    for k in range(3):
        assert np.array_equal(_np(fout["E"][k]), cpu["E"][k])
        assert np.array_equal(_np(fout["H"][k]), cpu["H"][k])
    g_eps = _np(bout["grad_eps_r"])
    if acc == "float32":
        assert np.array_equal(g_eps, cpu["g_eps"])
    else:
        assert h.rel_l2(g_eps, cpu["g_eps"]) < 1e-5
    # the user-facing gradient is the extended gradient folded onto the physical edges
    ext = torch.from_numpy(g_eps)
    pml = 10
    folded = ext.clone()
    for axis in range(2):
        size = folded.shape[axis] - 2 * pml
        core = folded.narrow(axis, pml, size).clone()
        core.select(axis, 0).add_(folded.narrow(axis, 0, pml).sum(dim=axis))
        core.select(axis, size - 1).add_(folded.narrow(axis, pml + size, pml).sum(dim=axis))
        folded = core
    # the fold sums replicated cells; CPU and GPU reductions order them differently
    np.testing.assert_allclose(eps.grad.cpu().numpy(), folded[:, :, 0].numpy(), rtol=2e-6, atol=1e-6)


# (11, 16, 90) and (12, 14, 96): rows long enough for the z-interior to run as its own launch
@pytest.mark.parametrize("shape,order", [((14, 12, 10), 4), ((12, 56, 44), 4), ((20, 40, 36), 8), ((16, 30, 26), 2),
                                         ((11, 16, 90), 4), ((12, 14, 96), 8)])
def test_compute_3d_matches_oracle(oracle_built, cuda_device, monkeypatch, shape, order):
    import deepgpr_amd as dg
    h = oracle_built
    calls = _capture(monkeypatch)
    gen = torch.Generator().manual_seed(7)
    eps = (1.0 + 8.0 * torch.rand(shape, generator=gen)).to(cuda_device).requires_grad_(True)
    sig = (0.01 * torch.rand(shape, generator=gen)).to(cuda_device).requires_grad_(True)
    from deepgpr_amd.preprocessing.wavelets import ricker
    nt = 60
    wave = ricker(8e8, nt, 1.2e-11, 1.5e-9).reshape(1, nt, 1).to(cuda_device).requires_grad_(True)
    src = torch.tensor([[[4, 5, 5]], [[9, 6, 4]]], device=cuda_device)
    rx = torch.tensor([[[3, 3, 3], [10, 8, 6]], [[5, 2, 7], [8, 9, 2]]], device=cuda_device)
    out = dg.compute(cuda_device, dx=0.02, dt=1.2e-11, source_amplitudes=wave, source_location=src,
                     receiver_location=rx, eps_r=eps, sigma=sig, pmlthick=4, fdtd_order=order, mode=3,
                     model_gradient_sampling_interval=3, gradient_accumulator="float32")
    (out[-1] ** 2).sum().backward()
    (_, fin, fout), (_, bin_, bout) = calls
    p = _problem(h, fin, bin_)
    cpu = h.run_cpu(p, "oracle")
    assert np.array_equal(_np(fout["receiver_data"]), cpu["rx"])
    assert np.array_equal(_np(bout["grad_eps_r"]), cpu["g_eps"])
    assert np.array_equal(_np(bout["grad_sigma"]), cpu["g_sig"])
    assert np.array_equal(_np(bout["grad_source"]), cpu["g_src"].reshape(_np(bout["grad_source"]).shape))
    assert wave.grad is not None and torch.isfinite(wave.grad).all()


def test_checkpoint_segments_equal_full_run(cuda_device):
    """Two chained segments with checkpoint_initial_field states equal one run
    (src: tests/test_cuda_fused_kernels.py:169, tests/test_external_pml.py:145)."""
    import deepgpr_amd as dg
    from deepgpr_amd import synthetic
    w = synthetic.c1_small_2d()
    nt = 300
    kw = dict(dx=w.dx, dt=w.dt, source_location=w.source_location.to(cuda_device),
              receiver_location=w.receiver_location.to(cuda_device), eps_r=w.eps_r.to(cuda_device),
              sigma=w.sigma.to(cuda_device), save_wavefield_history=False)
    wave = w.source_amplitudes[:, :nt].to(cuda_device)
    full = dg.compute(cuda_device, source_amplitudes=wave, **kw)[-1]
    E, H, P = dg.checkpoint_initial_field(device=cuda_device, dx=w.dx, dt=w.dt, source_amplitudes=wave,
                                          source_location=kw["source_location"],
                                          receiver_location=kw["receiver_location"], er=kw["eps_r"],
                                          se=kw["sigma"])
    half = nt // 2
    o1 = dg.compute(cuda_device, source_amplitudes=wave[:, :half], E=E, H=H, PML=P, **kw)
    o2 = dg.compute(cuda_device, source_amplitudes=wave[:, half:], E=o1[1], H=o1[2], PML=o1[3], **kw)
    seg = torch.cat([o1[-1], o2[-1]], dim=1)
    assert torch.equal(seg, full)
