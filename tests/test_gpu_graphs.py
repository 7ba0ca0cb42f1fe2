"""Forward-only 2D TM solves replay their step loop as a captured graph (32
steps per replay, the step read from a device counter, the remainder
launched one by one). The result must equal the same solve without the
graph: a forward that also saves its history takes the one-launch-per-step
loop and is itself checked against the oracle elsewhere."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(dev, nt, order):
    from deepgpr_amd.preprocessing.wavelets import ricker
    gen = torch.Generator().manual_seed(order)
    nx, ny = 70, 90
    eps = (1.0 + 8.0 * torch.rand((nx, ny), generator=gen)).to(dev)
    sig = (0.01 * torch.rand((nx, ny), generator=gen)).to(dev)
    dt = 2.0e-11
    wave = ricker(4e8, nt, dt, 2.5e-9).reshape(1, nt, 1).to(dev)
    src = torch.tensor([[[20, 30, 0]], [[41, 62, 0]]], device=dev)
    rx = torch.tensor([[[10, 10, 0], [35, 50, 0], [60, 85, 0]], [[5, 70, 0], [40, 44, 0], [66, 3, 0]]], device=dev)
    return dict(dx=0.02, dt=dt, source_amplitudes=wave, source_location=src, receiver_location=rx, eps_r=eps,
                sigma=sig, pmlthick=8, fdtd_order=order, mode=2)


@pytest.mark.parametrize("order,nt", [(2, 70), (4, 96), (8, 131)])
def test_graphed_forward_equals_step_loop(cuda_device, order, nt):
    import deepgpr_amd as dg
    kw = _inputs(cuda_device, nt, order)
    print(f"[graphs] order={order} nt={nt}", flush=True)
    with torch.no_grad():
        ref = dg.compute(cuda_device, save_wavefield_history=True, **kw)
        torch.cuda.synchronize()
        for rep in range(3):
            out = dg.compute(cuda_device, save_wavefield_history=False, **kw)
            # eager work and a synchronisation between the graphed solves
            scratch = torch.randn(1 << 16, device=cuda_device).cumsum(0)
            torch.cuda.synchronize()
            assert torch.isfinite(scratch).all()
            assert torch.equal(out[-1], ref[-1]), rep
            for a, b in zip(out[1] + out[2], ref[1] + ref[2]):
                assert torch.equal(a, b), rep
    assert ref[-1].abs().max() > 0
