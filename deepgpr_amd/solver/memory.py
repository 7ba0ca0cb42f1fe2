"""Device-memory footprint of a solve (src/DeepGPR/solver/memory.py:30-268),
used for the print_parameters preview and for admission by HBM footprint."""
from __future__ import annotations

import math
import os
from typing import Dict

from .pml import pml_phi_element_count
from .storage import int8_history_layout, saved_history_shape, saved_time_steps

_BYTES = {"float32": 4, "float16": 2, "bfloat16": 2}


def estimate_compute_memory(nx: int, ny: int, nz: int, nt: int, nstep: int, pml, *, mode: int = 3,
                            sampling_interval: int = 1, storage: str = "float32", compression: str = "none",
                            block_size=(4, 4, 4), reconstruct_rhs: bool = False, gradient: bool = True,
                            fp64_accumulators: bool = True, nrx: int = 1) -> Dict[str, int]:
    """Bytes per category for an extended grid (nx, ny, nz) with the Yee halo."""
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
        if nz > 1 and os.environ.get("DEEPGPR_AMD_3D_SWEEP") != "xmarch":
            # the one-thread-per-cell 3D adjoint keeps six channel weights per
            # cell for each half step; the x-marching sweeps form them on the fly
            out["adjoint_weights"] = 12 * nstep * (nx + 1) * (ny + 1) * (nz + 1) * 4
    out["total"] = sum(out.values())
    return out


def format_compute_preview(config) -> str:
    est = estimate_compute_memory(
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
