"""The E-only history (R^n rebuilt from consecutive E frames) against the
stored R history through compute() (restates the reference's
tests/test_rhs_history.py:27-101 and 190-233 on the HIP device)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(dev, mode, nt=60):
    import deepgpr_amd as dg
    if mode == 2:
        nx, ny = 18, 21
        x = torch.linspace(0.0, 1.0, nx)[:, None]
        y = torch.linspace(0.0, 1.0, ny)[None, :]
        eps_r = 4.0 + 0.3 * x + 0.2 * y
        sigma = 2.0e-4 + 1.0e-4 * x + 0.0 * y
        src = torch.tensor([[[8, 7, 0]], [[10, 9, 0]]], dtype=torch.int32)
        rx = torch.tensor([[[7, 12, 0], [9, 14, 0]], [[6, 13, 0], [11, 15, 0]]], dtype=torch.int32)
        spacing, dt, pml = (0.020, 0.016, 0.012), 2.5e-11, 3
    else:
        shape = (12, 13, 11)
        x = torch.linspace(0.0, 1.0, shape[0])[:, None, None]
        eps_r = 4.0 + 0.3 * x + torch.zeros(shape)
        sigma = 2.0e-4 + 1.0e-4 * x + torch.zeros(shape)
        src = torch.tensor([[[6, 6, 5]]], dtype=torch.int32)
        rx = torch.tensor([[[5, 8, 6], [7, 4, 5]]], dtype=torch.int32)
        spacing, dt, pml = (0.020, 0.018, 0.016), 1.5e-11, [2, 1, 2, 3, 1, 2]
    source = dg.ricker(4.0e8, nt, dt, 1.5e-9).reshape(1, nt, 1)
    common = dict(device=dev, dx=spacing, dt=dt, source_amplitudes=source.to(dev), source_location=src.to(dev),
                  receiver_location=rx.to(dev), pmlthick=pml, fdtd_order=4, mode=mode)
    return eps_r.to(dev), sigma.to(dev), common


def _run(dev, mode, rhs, mutate=None):
    import deepgpr_amd as dg
    eps_base, sigma_base, common = _case(dev, mode)
    eps_r = eps_base.clone().requires_grad_(True)
    sigma = sigma_base.clone().requires_grad_(True)
    source = common.pop("source_amplitudes").clone().requires_grad_(True)
    result = dg.compute(eps_r=eps_r, sigma=sigma, source_amplitudes=source, wavefield_rhs_history=rhs, **common)
    node = result[-1].grad_fn
    histories = [getattr(node, n) for n in ("E_saved", "R_saved", "E_final") if getattr(node, n, None) is not None]
    history_bytes = sum(t.numel() * t.element_size() for t in histories)
    data = result[-1]
    if mutate is not None:
        mutate(result)
    (data - 0.9 * data.detach()).square().sum().backward()
    return {"data": data.detach(), "history": result[0].detach(), "eps": eps_r.grad, "sigma": sigma.grad,
            "source": source.grad, "history_bytes": history_bytes}


@pytest.mark.parametrize("mode", [2, 3])
def test_float32_gradients_are_bitwise_identical(cuda_device, mode):
    print(f"[gpu rhs history] stored vs rebuilt, mode {mode}", flush=True)
    stored = _run(cuda_device, mode, "stored")
    rebuilt = _run(cuda_device, mode, "reconstructed")
    for key in ("data", "history", "eps", "sigma", "source"):
        assert torch.equal(stored[key], rebuilt[key]), key
    frame_bytes = stored["history_bytes"] // 2 // 60  # nt 60 at sampling interval 1
    assert rebuilt["history_bytes"] == stored["history_bytes"] // 2 + frame_bytes


def test_stored_rhs_equals_rebuilt_rhs_bitwise(cuda_device):
    import deepgpr_amd as dg
    print("[gpu rhs history] stored R against (E^(n+1) - ca E^n) / cb", flush=True)
    eps_r, sigma, common = _case(cuda_device, 2)
    eps_r = eps_r.clone().requires_grad_(True)
    result = dg.compute(eps_r=eps_r, sigma=sigma, wavefield_rhs_history="stored", **common)
    node = result[-1].grad_fn
    e_saved, r_saved = node.E_saved, node.R_saved
    final = result[1][2][:, :-1, :-1, :-1]
    later = torch.cat((e_saved[1:], final.unsqueeze(0)), dim=0)
    # ca and cb in float32 with the update-coefficient operations
    eps_pad = node.saved_tensors[15][:-1, :-1, :-1]
    sigma_pad = node.saved_tensors[16][:-1, :-1, :-1]
    e_term = torch.tensor(dg.config.constants.VACUUM_PERMITTIVITY, dtype=torch.float32, device=cuda_device)
    e_term = e_term * eps_pad / torch.tensor(common["dt"], dtype=torch.float32, device=cuda_device)
    s_term = 0.5 * sigma_pad
    ca = (e_term - s_term) / (e_term + s_term)
    cb = 1.0 / (e_term + s_term)
    assert torch.equal((later - ca * e_saved) / cb, r_saved)


def test_final_frame_is_independent_of_returned_states(cuda_device):
    print("[gpu rhs history] final frame and returned states", flush=True)

    def advance_returned_states(result):
        with torch.no_grad():
            for t in (*result[1], *result[2]):
                t.mul_(-3.0).add_(1.0)

    reference = _run(cuda_device, 2, "reconstructed")
    mutated = _run(cuda_device, 2, "reconstructed", mutate=advance_returned_states)
    for key in ("eps", "sigma", "source"):
        assert torch.equal(reference[key], mutated[key]), key
