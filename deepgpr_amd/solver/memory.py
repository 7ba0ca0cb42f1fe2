"""Device-memory footprint of a solve: ``estimate_compute_memory`` with the
reference's keywords and result keys (src/DeepGPR/solver/memory.py:30-246),
and ``solve_footprint``, the compact byte model behind the print_parameters
preview, admission by HBM footprint and the segment planner."""
from __future__ import annotations

import math
import os
from typing import Any, Dict, Optional, Sequence

import torch

from ..config.defaults import DEFAULT_GRADIENT_ACCUMULATOR
from ..utils.exceptions import NativeLibraryError
from .pml import pml_phi_element_count
from .storage import (default_compression_block_size, history_spatial_shape, int8_history_layout,
                      saved_history_shape, saved_time_steps)

_BYTES = {"float32": 4, "float16": 2, "bfloat16": 2}

# recommended capacities are the peaks with this margin, as in the reference
MEMORY_SAFETY_MARGIN = 1.20


def _adjoint_weight_cells(nz: int) -> bool:
    # the one-thread-per-cell 3D adjoint keeps six FP32 channel weights per
    # cell and shot for each half step; the x-marching sweeps form them on
    # the fly, and the 2D Ez-TM kernels recompute them on a halo
    return nz > 1 and os.environ.get("DEEPGPR_AMD_3D_SWEEP") != "xmarch"


def estimate_compute_memory(*, device, nx: int, ny: int, nz: int, nt: int, nstep: int, nsr: int, nrx: int,
                            source_waveforms: int, pml: Sequence[int], mode: int, sampling_interval: int,
                            storage_dtype: torch.dtype, use_async_offload: bool, er_requires_grad: bool,
                            se_requires_grad: bool, save_wavefield_history: bool = True,
                            wavefield_compression: str = "none", compression_block_size=None,
                            reconstruct_rhs: Optional[bool] = None, history_region: str = "extended",
                            gradient_accumulator: str = DEFAULT_GRADIENT_ACCUMULATOR) -> Dict[str, Any]:
    """Tensor payload of one compute() call on a HIP device, in bytes per
    category plus peaks (the reference's keywords and result keys; nx, ny, nz
    are the extended grid with the CPML). The figures are this build's:
    histories stay in HBM (use_async_offload changes nothing and
    effective_async_offload is False); a 3D adjoint keeps twelve FP32 channel
    weights per cell and shot in library scratch (adjoint_stencil_weights;
    2D models are priced for the Ez-TM kernels of the default Ez source,
    which keep none);
    FP64 accumulators (gradient_accumulator="float64", the default) add 16 B
    per model cell to material_gradients during the backward; no exact FP32
    frame is kept beside a low-precision history. The library has no CPU
    backend: a CPU device raises NativeLibraryError, as compute() does."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise NativeLibraryError(f"device {dev} has no backend in this build; the estimate is for HIP devices")
    f32 = 4
    storage_bytes = torch.empty(0, dtype=storage_dtype).element_size()
    model_cells = nx * ny * nz
    field_cells = (nx + 1) * (ny + 1) * (nz + 1)
    int8 = wavefield_compression == "int8"
    block = (tuple(compression_block_size) if compression_block_size is not None
             else default_compression_block_size(2 if nz == 1 else 3)) if int8 else None
    frame_cells = tuple(int(v) for v in history_spatial_shape(nx, ny, nz, pml, history_region, block))
    components = 3 if mode == 3 else 1
    nt_saved = saved_time_steps(nt, sampling_interval)
    needs_backward = bool(save_wavefield_history and (er_requires_grad or se_requires_grad))
    if reconstruct_rhs is None:  # compute()'s "auto": E-only for lossless FP32 at every step
        reconstruct_rhs = sampling_interval == 1 and not int8 and storage_dtype == torch.float32
    e_only = bool(reconstruct_rhs and needs_backward)

    def history_bytes(frames: int) -> int:
        shape = saved_history_shape(mode, frames, nstep, *frame_cells)
        if int8:
            return int8_history_layout(shape, block)["packed_bytes"]
        return math.prod(shape) * storage_bytes

    one = history_bytes(nt_saved) if save_wavefield_history else 0
    final = history_bytes(1) if save_wavefield_history and e_only else 0
    saved = one + final if e_only else (1 + int(needs_backward)) * one

    model = (2 * model_cells + 3 * field_cells) * f32
    fields = 6 * nstep * field_cells * f32
    psi = pml_phi_element_count(nx, ny, nz, nstep, pml) * f32
    cpml_coefficients = 8 * sum(int(p) for p in pml) * f32 + 7 * sum(int(p) > 0 for p in pml) * 4
    acquisition = source_waveforms * nt * f32 + nstep * (nsr + nrx) * 3 * 4
    core = model + fields + psi + cpml_coefficients + acquisition
    coefficients = 6 * field_cells * f32
    receivers = nstep * nt * nrx * f32
    gradients = (int(er_requires_grad) + int(se_requires_grad)) * model_cells * f32
    if needs_backward and gradient_accumulator == "float64":
        gradients += 2 * model_cells * 8
    adjoint = fields + psi if needs_backward else 0
    receiver_adjoint = receivers if needs_backward else 0
    weights = 12 * nstep * field_cells * f32 if needs_backward and _adjoint_weight_cells(nz) else 0
    forward_peak = core + saved + coefficients + receivers
    backward_peak = forward_peak + gradients + adjoint + receiver_adjoint + weights
    peak = max(forward_peak, backward_peak)
    return {
        "model_and_padded_materials": model,
        "electric_and_magnetic_fields": fields,
        "cpml_auxiliary_fields": psi,
        "cpml_coefficients": cpml_coefficients,
        "source_and_locations": acquisition,
        "saved_gradient_wavefields": saved,
        "final_e_frame": final,
        "packed_history_per_quantity": one if int8 else 0,
        "fdtd_update_coefficients": coefficients,
        "receiver_data": receivers,
        "material_gradients": gradients,
        "adjoint_fields_and_cpml": adjoint,
        "receiver_adjoint": receiver_adjoint,
        "adjoint_stencil_weights": weights,
        "low_precision_exact_snapshot": 0,
        "cuda_transfer_buffers": 0,
        "estimated_device_peak": peak,
        "estimated_host_peak": 0,
        "recommended_device_capacity": int(peak * MEMORY_SAFETY_MARGIN),
        "recommended_host_capacity": 0,
        "nt_saved": nt_saved,
        "components_saved": components,
        "history_shape": frame_cells,
        "reconstruct_rhs": e_only,
        "effective_async_offload": False,
    }


def solve_footprint(nx: int, ny: int, nz: int, nt: int, nstep: int, pml, *, mode: int = 3,
                    sampling_interval: int = 1, storage: str = "float32", compression: str = "none",
                    block_size=(4, 4, 4), reconstruct_rhs: bool = False, gradient: bool = True,
                    fp64_accumulators: bool = True, nrx: int = 1) -> Dict[str, int]:
    """Bytes per category for an extended grid (nx, ny, nz) with the Yee halo,
    and their "total": the byte model admission and the segment planner use
    (extended-grid history frames, so it never underestimates)."""
    field = nstep * (nx + 1) * (ny + 1) * (nz + 1) * 4
    cells = nx * ny * nz
    nt_saved = saved_time_steps(nt, sampling_interval)
    shape = saved_history_shape(mode, nt_saved, nstep, nx, ny, nz)
    if compression == "int8":
        one = int8_history_layout(shape, tuple(block_size))["packed_bytes"]
        final = int8_history_layout(saved_history_shape(mode, 1, nstep, nx, ny, nz), tuple(block_size))["packed_bytes"]
    else:
        one = math.prod(shape) * _BYTES[storage]
        final = math.prod(saved_history_shape(mode, 1, nstep, nx, ny, nz)) * _BYTES[storage]
    history = one if not gradient else (one + (final if reconstruct_rhs else one))
    out = {
        "fields": 6 * field,
        "cpml": 4 * pml_phi_element_count(nx, ny, nz, nstep, pml),
        "materials": 3 * (nx + 1) * (ny + 1) * (nz + 1) * 4 + 6 * (nx + 1) * (ny + 1) * (nz + 1) * 4,
        "history": history,
        "receivers": nstep * nt * nrx * 4,
    }
    if gradient:
        out["adjoint_fields"] = 6 * field
        out["adjoint_cpml"] = 2 * out["cpml"]
        out["gradients"] = 2 * cells * 4 + (2 * cells * 8 if fp64_accumulators else 0)
        if _adjoint_weight_cells(nz):
            out["adjoint_weights"] = 12 * nstep * (nx + 1) * (ny + 1) * (nz + 1) * 4
    out["total"] = sum(out.values())
    return out


# report rows: (label, estimate key), the rows of the reference's report
# (src/DeepGPR/solver/memory.py:249-264) with this build's scratch terms
_REPORT_ROWS = (
    ("model and padded materials", "model_and_padded_materials"),
    ("electric and magnetic fields", "electric_and_magnetic_fields"),
    ("CPML auxiliary fields", "cpml_auxiliary_fields"),
    ("CPML coefficients", "cpml_coefficients"),
    ("source and acquisition locations", "source_and_locations"),
    ("saved E_saved and R_saved wavefields", "saved_gradient_wavefields"),
    ("  of which final E frame (E-only history)", "final_e_frame"),
    ("FDTD update coefficients", "fdtd_update_coefficients"),
    ("receiver data", "receiver_data"),
    ("material gradients (with accumulators)", "material_gradients"),
    ("adjoint fields and CPML", "adjoint_fields_and_cpml"),
    ("receiver adjoint", "receiver_adjoint"),
    ("3D adjoint channel weights", "adjoint_stencil_weights"),
)


def _mib(nbytes: int) -> str:
    return f"{nbytes / 2**20:.2f} MiB"


def format_compute_preview(config, *, source_waveforms: int, er_requires_grad: bool, se_requires_grad: bool,
                           save_directory=None) -> str:
    """The print_parameters report of compute() (in the manner of
    src/DeepGPR/solver/memory.py:267-456): the resolved options, then the
    device-memory estimate of estimate_compute_memory by category."""
    st = config.storage
    est = estimate_compute_memory(
        device=config.device, nx=config.nx, ny=config.ny, nz=config.nz, nt=config.nt, nstep=config.nstep,
        nsr=config.nsr, nrx=config.nrx, source_waveforms=source_waveforms, pml=config.pml, mode=config.mode,
        sampling_interval=config.sampling_interval, storage_dtype=st.dtype,
        use_async_offload=config.use_async_offload, er_requires_grad=er_requires_grad,
        se_requires_grad=se_requires_grad, save_wavefield_history=config.save_wavefield_history,
        wavefield_compression=st.compression,
        compression_block_size=st.block_size if st.compression == "int8" else None,
        reconstruct_rhs=config.reconstruct_rhs, history_region=config.history_region,
        gradient_accumulator=config.grad_accumulator)
    dx, dy, dz = config.grid_spacing
    lines = [
        "=== DeepGPR compute preview ===",
        f"device: {config.device}",
        f"dx / dy / dz: {dx:g} / {dy:g} / {dz:g} m, dt: {config.dt:g} s",
        f"extended grid (nx, ny, nz): ({config.nx}, {config.ny}, {config.nz}), CPML faces: {config.pml}",
        f"time steps: {config.nt}, shots: {config.nstep}, sources / receivers per shot: {config.nsr} / {config.nrx}",
        f"FDTD order / gradient mode: {config.fdtd_order} / {config.mode}",
        f"source / receiver component: {config.source_direction} / {config.receiver_component}",
        f"2D Ez-TM kernels: {'yes' if config.tm2d_fast_path else 'no'}",
        f"history: save={config.save_wavefield_history}, storage={st.dtype}, compression={st.compression}"
        + (f" (block {tuple(st.block_size)})" if st.compression == "int8" else "")
        + f", sampling interval={config.sampling_interval}, region={config.history_region}",
        f"R history: {'rebuilt from E (E-only history)' if est['reconstruct_rhs'] else 'stored'}",
        f"gradient accumulator: {config.grad_accumulator}",
        f"async offload: requested={config.use_async_offload}, effective=False (histories stay in device memory)",
        f"forward wavefield save directory: {save_directory if save_directory is not None else 'disabled'}",
        "memory estimate (tensor payload):",
    ]
    lines += [f"  {label}: {_mib(est[key])}" for label, key in _REPORT_ROWS]
    lines += [
        f"estimated peak HIP device memory: {_mib(est['estimated_device_peak'])}",
        f"recommended device capacity with 20% margin: {_mib(est['recommended_device_capacity'])}",
        "=== End DeepGPR compute preview ===",
    ]
    return "\n".join(lines)
