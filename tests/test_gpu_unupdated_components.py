"""Material derivatives at Yee components whose base E update is inactive
(restates the reference's tests/test_unupdated_e_material_gradients.py on the
HIP device, on the 2D Ez-TM kernels and on the general kernels).

One step from a unit initial field at such a component: E^1 = E^0 + cb *
correction, where a source adds -w/dx^2 and a selected CPML auxiliary +0.5.
The retained E^0 does not depend on the material, so d(data)/d(eps) =
correction * (-eps0/dt) * cb^2 and d(data)/d(sigma) = correction * (-1/2) *
cb^2: the adjoint and central differences must both give these closed
forms."""
import contextlib

import pytest
import torch

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def _path(path):
    from deepgpr_amd.solver import modeling
    original = modeling._tm2d_applies
    if path == "general":
        modeling._tm2d_applies = lambda *args, **kwargs: False
    try:
        yield
    finally:
        modeling._tm2d_applies = original


def _case(dev, path="fast", component=2, dims=2, rhs="reconstructed", source_value=0.0, cpml_correction=False,
          compression=None, storage_dtype=torch.float32, conversion_backend="auto"):
    import deepgpr_amd as dg
    shape = (12, 12) if dims == 2 else (12, 12, 8)
    # Ex is inactive at y = 0; Ey and Ez at x = 0
    cell = (4, 0, 3) if component == 0 else (0, 4, 3)
    if dims == 2 or cpml_correction:
        cell = (cell[0], cell[1], 0)
    pml = [0, 0, 0, 0, 2, 0] if cpml_correction else 0
    eps = torch.full(shape, 4.0, dtype=torch.float32, device=dev)
    sigma = torch.full(shape, 0.01, dtype=torch.float32, device=dev)
    source = torch.full((1, 1, 1), source_value, dtype=torch.float32, device=dev)
    location = torch.tensor([[cell]], dtype=torch.int32, device=dev)
    dt, dx = 2.5e-11, 0.02
    initial = dg.checkpoint_initial_field(device=dev, dx=dx, dt=dt, source_amplitudes=source,
                                          source_location=location, receiver_location=location, er=eps, se=sigma,
                                          pmlthick=pml, fdtd_order=2)
    states = [s.detach().clone() for group in initial for s in group]
    extended_cell = (cell[0], cell[1], cell[2] + (2 if cpml_correction else 0))
    states[component][(0, *extended_cell)] = 1.0
    if cpml_correction:
        # the z0 face's electric auxiliary that feeds Ey; its q = 0 profile
        # values are RA = RB = RE = 1, RF = 0
        states[6 + 4 * 4 + 1][0, 0, 4, 0] = 0.5
    options = dict(device=dev, dx=dx, dt=dt, source_amplitudes=source, source_location=location,
                   receiver_location=location, pmlthick=pml, fdtd_order=2, mode=2 if dims == 2 else 3,
                   source_direction=component, receiver_component=component, model_gradient_sampling_interval=1,
                   wavefield_storage_dtype=storage_dtype, wavefield_rhs_history=rhs,
                   wavefield_conversion_backend=conversion_backend)
    if compression:
        options.update(wavefield_compression=compression,
                       wavefield_compression_block_size=(4, 4) if dims == 2 else (4, 4, 1))

    def objective(er, se, backward=False):
        owned = [s.detach().clone() for s in states]
        with _path(path):
            result = dg.compute(eps_r=er, sigma=se, E=tuple(owned[:3]), H=tuple(owned[3:6]), PML=tuple(owned[6:]),
                                **options)
        data = result[-1]
        fast = bool(data.grad_fn.config.tm2d_fast_path) if data.grad_fn else None
        loss = 0.5 * data.double().square().sum()
        if backward:
            loss.backward()
        return float(loss.detach()), float(data.detach().flatten()[0]), fast

    er = eps.detach().clone().requires_grad_(True)
    se = sigma.detach().clone().requires_grad_(True)
    _, data, fast = objective(er, se, backward=True)
    physical_cell = cell[:dims]
    fd = {}
    for parameter, baseline, h in (("eps", eps, 0.04), ("sigma", sigma, 0.003)):
        direction = torch.zeros_like(baseline)
        direction[physical_cell] = h
        if parameter == "eps":
            plus, minus = objective(eps + direction, sigma)[0], objective(eps - direction, sigma)[0]
        else:
            plus, minus = objective(eps, sigma + direction)[0], objective(eps, sigma - direction)[0]
        fd[parameter] = (plus - minus) / (2 * h)
    correction = (0.5 if cpml_correction else 0.0) - source_value / dx ** 2
    eps0 = dg.config.constants.VACUUM_PERMITTIVITY
    cb = 1.0 / (eps0 * 4.0 / dt + 0.5 * 0.01)
    return dict(path=path, compression=compression, storage_dtype=storage_dtype, fast=fast,
                ad_eps=float(er.grad[physical_cell]), ad_sigma=float(se.grad[physical_cell]),
                fd_eps=fd["eps"], fd_sigma=fd["sigma"],
                analytic_eps=data * correction * (-eps0 / dt) * cb ** 2,
                analytic_sigma=data * correction * (-0.5) * cb ** 2)


def _check(row):
    lossy = row["compression"] or row["storage_dtype"] != torch.float32
    relative = 0.008 if lossy else 0.0005
    for parameter in ("eps", "sigma"):
        expected = row[f"analytic_{parameter}"]
        assert abs(row[f"ad_{parameter}"] - expected) <= max(2.0e-6, abs(expected) * relative), row
        assert abs(row[f"fd_{parameter}"] - expected) <= max(2.0e-5, abs(expected) * 0.001), row
    assert row["fast"] == (row["path"] == "fast"), row


@pytest.mark.parametrize("path,dims,component", [("fast", 2, 2), ("general", 2, 2), ("general", 3, 0),
                                                 ("general", 3, 1), ("general", 3, 2)])
def test_unupdated_components_and_boundary_sources(cuda_device, path, dims, component):
    print(f"[gpu unupdated] {path} {dims}D component {component}", flush=True)
    for rhs in ("stored", "reconstructed"):
        for source in (0.0, 0.0002):
            _check(_case(cuda_device, path, component, dims, rhs, source))


def test_unupdated_component_keeps_the_cpml_correction(cuda_device):
    print("[gpu unupdated] CPML correction", flush=True)
    for rhs in ("stored", "reconstructed"):
        _check(_case(cuda_device, "general", 1, 3, rhs, cpml_correction=True))


@pytest.mark.parametrize("path", ["fast", "general"])
def test_int8_unupdated_component_and_boundary_source(cuda_device, path):
    print(f"[gpu unupdated] INT8 {path}", flush=True)
    for rhs in ("stored", "reconstructed"):
        for source in (0.0, 0.0002):
            _check(_case(cuda_device, path, 2, 2, rhs, source, compression="int8"))


@pytest.mark.parametrize("path", ["fast", "general"])
def test_lossy_rhs_history_keeps_the_actual_component_coefficient(cuda_device, path):
    print(f"[gpu unupdated] FP16/BF16 {path}", flush=True)
    for backend in ("legacy", "native_scalar", "native_vec2"):
        for dtype in (torch.float16, torch.bfloat16):
            for source in (0.0, 0.0002):
                _check(_case(cuda_device, path, rhs="stored", source_value=source, storage_dtype=dtype,
                             conversion_backend=backend))
