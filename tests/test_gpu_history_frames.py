"""What the history holds, through compute() on the HIP device (restates the
reference's tests/test_cuda_fused_kernels.py): saved frame k is the E field
after k*S steps, in the extended or the physical box, with or without the
offload keyword; the FP16/BF16 conversion backends write the same bits; the
forward fields do not depend on the history format; and a forward split in
two at any step reproduces the whole run bit for bit."""
import warnings

import pytest
import torch

pytestmark = pytest.mark.gpu

STORAGE = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}


def _case(dev, mode, nt):
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
    # an early peak, so that the CPML cells carry nonzero fields within a few steps
    source = dg.ricker(4.0e8, nt, dt, 5.0e-10).reshape(1, nt, 1)
    common = dict(device=dev, dx=spacing, dt=dt, source_location=src.to(dev), receiver_location=rx.to(dev),
                  pmlthick=pml, fdtd_order=4, mode=mode)
    return eps_r.to(dev), sigma.to(dev), source.to(dev), common


def _history(dev, mode, nt, material_grad=True, **options):
    import deepgpr_amd as dg
    eps_r, sigma, source, common = _case(dev, mode, nt)
    eps_r = eps_r.clone().requires_grad_(material_grad)
    source = source.clone().requires_grad_(True)
    result = dg.compute(eps_r=eps_r, sigma=sigma, source_amplitudes=source, **common, **options)
    return result[-1].grad_fn.E_saved.detach(), [t.detach() for t in result[1]], result[-1].detach()


def _fields_after(dev, mode, steps):
    import deepgpr_amd as dg
    eps_r, sigma, source, common = _case(dev, mode, steps)
    result = dg.compute(eps_r=eps_r, sigma=sigma, source_amplitudes=source, save_wavefield_history=False, **common)
    return [t.detach() for t in result[1]]


def _frame(fields, mode, origin, extent):
    (x0, y0, z0), (hx, hy, hz) = origin, extent
    crops = [f[:, x0:x0 + hx, y0:y0 + hy, z0:z0 + hz] for f in fields]
    return torch.stack(crops) if mode == 3 else crops[2]


@pytest.mark.parametrize("mode", [2, 3])
def test_saved_frames_are_the_fields_after_each_step(cuda_device, mode):
    from deepgpr_amd.solver.storage import history_box
    print(f"[gpu history frames] frames, mode {mode}", flush=True)
    nt = 30
    _, _, _, common = _case(cuda_device, mode, nt)
    reference = {steps: _fields_after(cuda_device, mode, steps) for steps in range(1, nt)}
    nx, ny, nz = (f - 1 for f in reference[1][0].shape[1:])
    for region in ("extended", "physical"):
        for offload in (False, True):
            for interval in (1, 3):
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore", RuntimeWarning)  # 3 exceeds this short source's bound
                    history, _, _ = _history(cuda_device, mode, nt, wavefield_history_region=region,
                                             use_async_offload=offload, model_gradient_sampling_interval=interval,
                                             material_grad=region != "physical")
                origin, extent = history_box(nx, ny, nz, common["pmlthick"], region)
                frames = history if mode == 2 else history.transpose(0, 1)
                case = (region, offload, interval)
                assert frames.shape[0] == (nt + interval - 1) // interval, case
                assert int(torch.count_nonzero(frames[0])) == 0, case
                if region == "extended":  # the low-x CPML slab is nonzero, so its updates are covered
                    assert int(torch.count_nonzero(frames[-1][..., :common["pmlthick"][0], :, :])) > 0, case
                for index in range(1, frames.shape[0]):
                    assert torch.equal(frames[index], _frame(reference[index * interval], mode, origin, extent)), \
                        (case, index)


@pytest.mark.parametrize("mode", [2, 3])
def test_low_precision_frames_do_not_depend_on_the_conversion_backend(cuda_device, mode):
    print(f"[gpu history frames] conversion backends, mode {mode}", flush=True)
    for storage in ("fp16", "bf16"):
        for offload in (False, True):
            runs = [_history(cuda_device, mode, 24, wavefield_storage_dtype=STORAGE[storage],
                             wavefield_conversion_backend=backend, use_async_offload=offload)
                    for backend in ("native_scalar", "native_vec2", "legacy")]
            for other in runs[1:]:
                assert torch.equal(runs[0][0], other[0]), (storage, offload)
                assert torch.equal(runs[0][2], other[2]), (storage, offload)


@pytest.mark.parametrize("mode", [2, 3])
def test_forward_does_not_depend_on_the_history_format(cuda_device, mode):
    print(f"[gpu history frames] history formats, mode {mode}", flush=True)
    options = (dict(wavefield_storage_dtype=torch.float32),
               dict(wavefield_storage_dtype=torch.float16, use_async_offload=True),
               dict(wavefield_compression="int8"),
               dict(wavefield_history_region="physical", wavefield_rhs_history="stored"))
    runs = [_history(cuda_device, mode, 24, material_grad=o.get("wavefield_history_region") != "physical", **o)
            for o in options]
    reference = _fields_after(cuda_device, mode, 24)
    for option, (_, fields, data) in zip(options, runs):
        assert torch.equal(data, runs[0][2]), option
        for field, expected in zip(fields, reference):
            assert torch.equal(field, expected), option


@pytest.mark.parametrize("mode", [2, 3])
def test_split_forward_reproduces_the_whole_run(cuda_device, mode):
    import deepgpr_amd as dg
    nt, split = 30, 13
    eps_r, sigma, source, common = _case(cuda_device, mode, nt)
    full = dg.compute(eps_r=eps_r, sigma=sigma, source_amplitudes=source, save_wavefield_history=False, **common)
    first = dg.compute(eps_r=eps_r, sigma=sigma, source_amplitudes=source[:, :split], save_wavefield_history=False,
                       **common)
    second = dg.compute(eps_r=eps_r, sigma=sigma, source_amplitudes=source[:, split:],
                        E=tuple(t.clone() for t in first[1]), H=tuple(t.clone() for t in first[2]),
                        PML=tuple(t.clone() for t in first[3]), save_wavefield_history=False, **common)
    assert torch.equal(torch.cat((first[-1], second[-1]), dim=1), full[-1])
    for part, whole in zip((*second[1], *second[2], *second[3]), (*full[1], *full[2], *full[3])):
        assert torch.equal(part, whole)
