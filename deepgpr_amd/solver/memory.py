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


def format_compute_preview(config) -> str:
    est = solve_footprint(
        config.nx, config.ny, config.nz, config.nt, config.nstep, config.pml, mode=config.mode,
        sampling_interval=config.sampling_interval, storage=str(config.storage.dtype).replace("torch.", ""),
        compression=config.storage.compression, block_size=config.storage.block_size or (4, 4, 4),
        reconstruct_rhs=config.reconstruct_rhs, fp64_accumulators=config.grad_accumulator == "float64",
        nrx=config.nrx)
    lines = [
        "deepgpr_amd compute",
        f"  device={config.device} order={config.fdtd_order} mode={config.mode}",
        f"  extended grid=({config.nx}, {config.ny}, {config.nz}) nt={config.nt} shots={config.nstep}",
        f"  history: {config.storage.dtype} compression={config.storage.compression} S={config.sampling_interval}"
        f" E-only={config.reconstruct_rhs} region={config.history_region}",
    ]
    lines += [f"  {k:>15s}: {v / 2**20:12.1f} MiB" for k, v in est.items()]
    return "\n".join(lines)
