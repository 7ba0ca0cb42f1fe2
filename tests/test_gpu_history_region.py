"""The physical-region history through compute() on the HIP device (restates
the reference's tests/test_history_region.py:28-114 and 173-300). The
reference bounds source gradients by 2e-6 on its GPU (atomic adjoint
scatter); this adjoint is a deterministic gather, so they are compared bit for
bit."""
import pytest
import torch

pytestmark = pytest.mark.gpu

STORAGE = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}


def _case(dev, mode, nt=40):
    import deepgpr_amd as dg
    if mode == 2:
        nx, ny = 17, 19
        x = torch.linspace(0.0, 1.0, nx)[:, None]
        y = torch.linspace(0.0, 1.0, ny)[None, :]
        eps_r = 4.0 + 0.3 * x + 0.2 * y
        sigma = 2.0e-4 + 1.0e-4 * x + 0.0 * y
        src = torch.tensor([[[7, 6, 0]], [[9, 8, 0]]], dtype=torch.int32)
        rx = torch.tensor([[[6, 11, 0], [8, 13, 0]], [[5, 12, 0], [10, 14, 0]]], dtype=torch.int32)
        spacing, dt, pml = (0.020, 0.016, 0.012), 2.5e-11, [3, 5, 4, 2, 0, 0]
    else:
        shape = (12, 13, 11)
        x = torch.linspace(0.0, 1.0, shape[0])[:, None, None]
        eps_r = 4.0 + 0.3 * x + torch.zeros(shape)
        sigma = 2.0e-4 + 1.0e-4 * x + torch.zeros(shape)
        src = torch.tensor([[[6, 6, 5]]], dtype=torch.int32)
        rx = torch.tensor([[[5, 8, 6], [7, 4, 5]]], dtype=torch.int32)
        spacing, dt, pml = (0.020, 0.018, 0.016), 1.5e-11, [2, 1, 3, 2, 1, 2]
    source = dg.ricker(4.0e8, nt, dt, 1.5e-9).reshape(1, nt, 1)
    common = dict(device=dev, dx=spacing, dt=dt, source_amplitudes=source.to(dev), source_location=src.to(dev),
                  receiver_location=rx.to(dev), pmlthick=pml, fdtd_order=4, mode=mode)
    return eps_r.to(dev), sigma.to(dev), common


def _run(dev, mode, region, storage="fp32", rhs="auto", compression="none", offload=False, material_grad=True,
         pml=None):
    import deepgpr_amd as dg
    eps_base, sigma_base, common = _case(dev, mode)
    if pml is not None:
        common["pmlthick"] = pml
    eps_r = eps_base.clone().requires_grad_(material_grad)
    sigma = sigma_base.clone().requires_grad_(material_grad)
    source = common.pop("source_amplitudes").clone().requires_grad_(True)
    result = dg.compute(eps_r=eps_r, sigma=sigma, source_amplitudes=source, wavefield_storage_dtype=STORAGE[storage],
                        wavefield_rhs_history=rhs, wavefield_compression=compression, use_async_offload=offload,
                        wavefield_history_region=region, **common)
    node = result[-1].grad_fn
    history_bytes = sum(t.numel() * t.element_size() for t in (node.E_saved, node.R_saved, node.E_final)
                        if t is not None)
    data = result[-1]
    (data - 0.9 * data.detach()).square().sum().backward()
    return {"data": data.detach(), "history": result[0].detach(), "eps": eps_r.grad, "sigma": sigma.grad,
            "source": source.grad, "history_bytes": history_bytes, "pml": common["pmlthick"]}


def _crop(history, pml):
    x0, xm, y0, ym, z0, zm = pml
    nx, ny, nz = history.shape[-3:]
    return history[..., x0:nx - xm, y0:ny - ym, z0:nz - zm]


def test_forward_only_accepts_a_physical_history_with_cpml(cuda_device):
    import deepgpr_amd as dg
    eps, sigma, common = _case(cuda_device, 2)
    result = dg.compute(eps_r=eps.requires_grad_(), sigma=sigma.requires_grad_(), wavefield_history_region="physical",
                        save_wavefield_history=False, **common)
    assert result[0].numel() == 0 and bool(torch.isfinite(result[-1]).all())


@pytest.mark.parametrize("mode", [2, 3])
def test_physical_history_preserves_initial_state_gradients(cuda_device, mode):
    import deepgpr_amd as dg
    print(f"[gpu history region] initial-state gradients, mode {mode}", flush=True)
    eps, sigma, common = _case(cuda_device, mode)
    initial_args = {k: v for k, v in common.items() if k != "mode"}
    groups = dg.checkpoint_initial_field(er=eps, se=sigma, **initial_args)
    gen = torch.Generator().manual_seed(1670 + mode)
    initial = [(torch.randn(t.shape, generator=gen) * 1.e-3).to(cuda_device) for group in groups for t in group]
    gradients = []
    for region in ("extended", "physical"):
        owned = [t.clone().requires_grad_() for t in initial]
        result = dg.compute(eps_r=eps, sigma=sigma, E=tuple(owned[:3]), H=tuple(owned[3:6]), PML=tuple(owned[6:]),
                            wavefield_history_region=region, **common)
        loss = (result[-1].double().square().sum()
                + sum(t.double().square().sum() for group in result[1:4] for t in group))
        loss.backward()
        gradients.append([t.grad for t in owned])
    for full, physical in zip(*gradients):
        assert torch.equal(full, physical)


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("parameter", ["eps", "sigma"])
def test_physical_history_rejects_cpml_material_gradients(cuda_device, mode, parameter):
    import deepgpr_amd as dg
    eps, sigma, common = _case(cuda_device, mode)
    (eps if parameter == "eps" else sigma).requires_grad_()
    with pytest.raises(ValueError, match="physical.*PML|PML.*physical"):
        dg.compute(eps_r=eps, sigma=sigma, wavefield_history_region="physical", **common)


@pytest.mark.parametrize("mode", [2, 3])
def test_physical_history_preserves_forward_and_source_gradients(cuda_device, mode):
    print(f"[gpu history region] forward and source gradients, mode {mode}", flush=True)
    for storage, rhs, offload in (("fp32", "auto", False), ("fp32", "stored", True), ("fp16", "auto", False),
                                  ("bf16", "reconstructed", True)):
        extended = _run(cuda_device, mode, "extended", storage, rhs, offload=offload, material_grad=False)
        physical = _run(cuda_device, mode, "physical", storage, rhs, offload=offload, material_grad=False)
        case = (storage, rhs, offload)
        assert torch.equal(extended["data"], physical["data"]), case
        assert torch.equal(extended["source"], physical["source"]), case
        assert torch.equal(_crop(extended["history"], extended["pml"]), physical["history"]), case
        assert (physical["history_bytes"] * extended["history"].numel()
                == extended["history_bytes"] * physical["history"].numel()), case
        assert physical["history"].numel() / extended["history"].numel() < 0.75, case


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("rhs", ["stored", "reconstructed"])
def test_physical_history_without_cpml_keeps_material_gradients(cuda_device, mode, rhs):
    extended = _run(cuda_device, mode, "extended", rhs=rhs, pml=0)
    physical = _run(cuda_device, mode, "physical", rhs=rhs, pml=0)
    for key in ("data", "history", "eps", "sigma", "source"):
        assert torch.equal(extended[key], physical[key]), key


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("rhs", ["stored", "reconstructed"])
def test_int8_tiles_follow_the_full_grid_tiling(cuda_device, mode, rhs):
    import deepgpr_amd as dg
    from deepgpr_amd.solver.storage import history_box, saved_history_shape
    print(f"[gpu history region] INT8 tiles, mode {mode} {rhs}", flush=True)
    extended = _run(cuda_device, mode, "extended", rhs=rhs, compression="int8", material_grad=False)
    physical = _run(cuda_device, mode, "physical", rhs=rhs, compression="int8", material_grad=False)
    _, _, common = _case("cpu", mode)
    nx, ny, nz = (25, 25, 1) if mode == 2 else (15, 18, 14)
    block = (8, 8, 1) if mode == 2 else (4, 4, 4)
    origin, extent = history_box(nx, ny, nz, common["pmlthick"], "physical", block)
    nstep, nt_saved = extended["data"].shape[:2]
    full = dg.decompress_wavefield_history(extended["history"], saved_history_shape(mode, nt_saved, nstep, nx, ny, nz),
                                           block)
    part = dg.decompress_wavefield_history(physical["history"], saved_history_shape(mode, nt_saved, nstep, *extent),
                                           block)
    (x0, y0, z0), (hx, hy, hz) = origin, extent
    assert torch.equal(full[..., x0:x0 + hx, y0:y0 + hy, z0:z0 + hz], part)
    assert physical["history_bytes"] < extended["history_bytes"]
    assert torch.equal(extended["source"], physical["source"])
