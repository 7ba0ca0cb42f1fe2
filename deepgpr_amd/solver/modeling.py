"""``compute``: FDTD forward modelling with autograd on MI355X
(src/DeepGPR/solver/modeling.py:261-671; same keywords, defaults, return
values and in-place state semantics)."""
from __future__ import annotations

import math
import os
import warnings
from dataclasses import dataclass
from datetime import datetime
from typing import Any, Optional, Tuple

import torch

from ..config.constants import (
    FIELD_COMPONENTS, GRADIENT_ACCUMULATORS, INT8_SPECIALISED_REDUCTION_VOLUME, SPEED_OF_LIGHT,
    SUPPORTED_FDTD_ORDERS, SUPPORTED_GRADIENT_MODES,
)
from ..config.defaults import (
    DEFAULT_CONVERSION_BACKEND, DEFAULT_FDTD_ORDER, DEFAULT_GRADIENT_ACCUMULATOR, DEFAULT_GRADIENT_MODE,
    DEFAULT_GRADIENT_SAMPLING_INTERVAL, DEFAULT_INT8_REDUCTION_BACKEND, DEFAULT_PML_THICKNESS,
    DEFAULT_RECEIVER_COMPONENT, DEFAULT_SOURCE_COMPONENT, DEFAULT_WAVEFIELD_COMPRESSION,
    DEFAULT_WAVEFIELD_HISTORY_REGION, DEFAULT_WAVEFIELD_RHS_HISTORY,
)
from ..native.loader import library_supports, require_library
from ..preprocessing.grid import normalize_grid_spacing
from ..preprocessing.model_setup import PreparedModel, initialization, shift_locations
from ..utils.devices import resolve_device
from ..utils.logger import get_logger
from .autograd import TM2D_INACTIVE_STATE_INDICES, DeepGPR, SolverConfig, SolverTensors
from .fields import create_or_separate
from .pml import build_pml_coeffs, build_pml_phi
from .sampling import AUTO_SAMPLING_INTERVAL, recommended_sampling_interval, source_max_frequency
from .storage import (
    WavefieldStorageConfig, history_box, normalize_compression_block_size, normalize_int8_reduction_backend,
    normalize_wavefield_compression, normalize_wavefield_conversion_backend, normalize_wavefield_history_region,
    normalize_wavefield_rhs_history, normalize_wavefield_storage_dtype, resolve_rhs_reconstruction,
    saved_history_shape, saved_time_steps,
)

_LOG = get_logger(__name__)


def _alias(name: str, value: Any, alias: str, alias_value: Any) -> Any:
    if value is not None and alias_value is not None:
        raise TypeError(f"Specify only one of {name} and its deprecated alias {alias}.")
    return alias_value if value is None else value


def _storage_options(device, storage_dtype, compression, conversion, reduction, block, rate,
                     save_history: bool, async_offload: bool, has_output: bool):
    storage_dtype = normalize_wavefield_storage_dtype(storage_dtype)
    compression = normalize_wavefield_compression(compression)
    conversion = normalize_wavefield_conversion_backend(conversion)
    reduction = normalize_int8_reduction_backend(reduction)
    if conversion == "auto":
        conversion = ("native_scalar" if device.type == "cuda" and compression == "none"
                      and storage_dtype == torch.float16 else "legacy")
    if compression == "zfp":
        raise NotImplementedError("wavefield_compression='zfp' has no fused decoder; use 'int8' or 'none'.")
    if rate is not None:
        raise ValueError("wavefield_compression_rate is only meaningful for the optional ZFP backend.")
    if not save_history and async_offload:
        raise ValueError("use_async_offload=True is incompatible with save_wavefield_history=False "
                         "because there is no history to offload.")
    if not save_history and has_output:
        raise ValueError("save_forward_wavefield_path requires save_wavefield_history=True.")
    if compression == "int8":
        if device.type != "cuda":
            raise ValueError("wavefield_compression='int8' is GPU-only and GPU-resident.")
        if storage_dtype != torch.float32:
            raise ValueError("wavefield_storage_dtype must remain float32 when wavefield_compression='int8'.")
        if async_offload:
            raise ValueError("use_async_offload=True is incompatible with GPU-resident INT8 wavefield compression.")
    elif block is not None:
        raise ValueError("wavefield_compression_block_size is only valid when wavefield_compression='int8'.")
    if compression != "int8" and reduction not in ("auto", "current"):
        raise ValueError("int8_reduction_backend is only selectable when wavefield_compression='int8'.")
    if conversion != "legacy":
        if device.type != "cuda":
            raise ValueError("Native wavefield conversion backends are GPU-only.")
        if compression != "none" or storage_dtype not in (torch.float16, torch.bfloat16):
            raise ValueError("wavefield_conversion_backend is only selectable for uncompressed "
                             "float16/bfloat16 histories.")
    return storage_dtype, compression, conversion, reduction


def _int8_block(compression: str, block, spatial_mode: int, reduction: str):
    resolved = normalize_compression_block_size(block, spatial_mode) if compression == "int8" else None
    if reduction == "auto":
        reduction = ("cub_block" if compression == "int8"
                     and math.prod(resolved) == INT8_SPECIALISED_REDUCTION_VOLUME else "current")  # type: ignore
    if compression == "int8" and reduction != "current" and math.prod(resolved) != INT8_SPECIALISED_REDUCTION_VOLUME:  # type: ignore
        raise ValueError("The cub_block and warp_shuffle INT8 reduction backends need a compression block of "
                         "exactly 64 voxels.")
    return resolved, reduction


def _sampling(requested, automatic: bool, waveforms, dt, report: bool) -> int:
    if not automatic and (requested == 1 or not report):
        return int(requested)
    f_max = source_max_frequency(waveforms, dt)
    bound = recommended_sampling_interval(waveforms, dt)
    if automatic:
        if report:
            _LOG.info("model_gradient_sampling_interval='auto' selected %d (f_max=%.4e Hz).", bound, f_max)
        return bound
    if requested > bound:
        warnings.warn(f"model_gradient_sampling_interval={requested} exceeds the recommended bound {bound} = "
                      f"floor(1 / (4 f_max dt)) for the source spectrum (f_max={f_max:.4e} Hz, dt={float(dt):.4e} s). "
                      "Above the bound the sampled gradient loses accuracy quickly; pass 'auto' or an interval "
                      "within it.", RuntimeWarning, stacklevel=3)
    else:
        _LOG.info("model_gradient_sampling_interval=%d samples the gradient in time, within the recommended bound "
                  "%d (f_max=%.4e Hz).", requested, bound, f_max)
    return int(requested)


def _tm2d_applies(device, spatial_mode: int, source_direction: int, states, user_states: bool) -> bool:
    """2D Ez-TM fast path: nz = 1, Ez source and zero Ex/Ey/Hz (+ their CPML)."""
    if device.type != "cuda" or spatial_mode != 2 or source_direction != 2:
        return False
    if not library_supports(require_library(device), "deepgpr_supports_tm2d_fast_path"):
        return False
    if not user_states:
        return True
    inactive = [states[k] for k in TM2D_INACTIVE_STATE_INDICES]
    return int(sum(t.detach().count_nonzero() for t in inactive if t.numel())) == 0


def wavefield_directory(value) -> Optional[str]:
    """``save_forward_wavefield_path`` as a directory path, or None. A value that
    is not path-like raises TypeError, an existing non-directory
    NotADirectoryError (src/DeepGPR/postprocessing/io.py:17-37)."""
    if value is None:
        return None
    try:
        path = os.path.expanduser(os.fspath(value))
    except TypeError as exc:
        raise TypeError("save_forward_wavefield_path must be a string, path-like object, or None.") from exc
    if os.path.exists(path) and not os.path.isdir(path):
        raise NotADirectoryError(f"save_forward_wavefield_path is not a directory: {path}")
    return path


def _claim_wavefield_file(directory: str, label: str) -> str:
    """Create ``forward_wavefield_<label>.pt``, or the first free
    ``forward_wavefield_<label>_NN.pt`` (NN = 01, 02, ...), with exclusive
    creation so that concurrent savers never share a file
    (src/DeepGPR/postprocessing/io.py:40-57)."""
    n = 0
    while True:
        name = f"forward_wavefield_{label}.pt" if n == 0 else f"forward_wavefield_{label}_{n:02d}.pt"
        path = os.path.join(directory, name)
        try:
            fd = os.open(path, os.O_CREAT | os.O_EXCL | os.O_WRONLY, 0o666)
        except FileExistsError:
            n += 1
            continue
        os.close(fd)
        return path


def save_forward_wavefield(e_saved: torch.Tensor, directory: str, when: datetime, metadata=None) -> str:
    """Write a CPU copy of the history to ``forward_wavefield_HH-MM[_NN].pt`` and
    return its absolute path: the bare tensor or, with metadata,
    ``{"wavefield": tensor, **metadata}``. A failed write removes the file
    (src/DeepGPR/postprocessing/io.py:60-92)."""
    os.makedirs(directory, exist_ok=True)
    path = _claim_wavefield_file(directory, when.strftime("%H-%M"))
    data = e_saved.detach().to(device="cpu")
    payload = data if metadata is None else {"wavefield": data, **metadata}
    try:
        torch.save(payload, path)
    except BaseException:
        if os.path.exists(path):
            os.remove(path)
        raise
    return os.path.realpath(path)


@dataclass
class PreparedSolve:
    """Inputs of one solve as ``compute`` hands them to the solver."""

    config: SolverConfig
    tensors: SolverTensors
    eps_ext: torch.Tensor
    sig_ext: torch.Tensor
    waveforms: torch.Tensor
    electric: tuple
    magnetic: tuple
    phi: tuple
    out_dir: Optional[str]
    when: Optional[datetime]
    mode: int
    region: str
    block: Optional[tuple]
    sampling_interval: int


def prepare_compute(device, dx=None, dt=None, source_amplitudes=None, source_location=None, receiver_location=None,
            er=None, se=None, mr=None, E=None, H=None, PML=None, pmlthick=DEFAULT_PML_THICKNESS,
            source_direction=DEFAULT_SOURCE_COMPONENT, reciever_direction=DEFAULT_RECEIVER_COMPONENT,
            model_gradient_sampling_interval=DEFAULT_GRADIENT_SAMPLING_INTERVAL,
            wavefield_storage_dtype=torch.float32, use_async_offload=False, fdtd_order=DEFAULT_FDTD_ORDER,
            mode=DEFAULT_GRADIENT_MODE, debug=False, print_parameters=False, save_forward_wavefield_path=None, *,
            eps_r=None, sigma=None, mu_r=None, receiver_component=None, save_wavefield_history=True,
            wavefield_compression=DEFAULT_WAVEFIELD_COMPRESSION, wavefield_compression_block_size=None,
            wavefield_conversion_backend=DEFAULT_CONVERSION_BACKEND,
            int8_reduction_backend=DEFAULT_INT8_REDUCTION_BACKEND, wavefield_compression_rate=None,
            wavefield_rhs_history=DEFAULT_WAVEFIELD_RHS_HISTORY,
            wavefield_history_region=DEFAULT_WAVEFIELD_HISTORY_REGION, pml_max_velocity=SPEED_OF_LIGHT,
            gradient_accumulator=DEFAULT_GRADIENT_ACCUMULATOR):
    """Everything ``compute`` does before the solve: validation, model
    extension, CPML coefficients and auxiliaries, states and the solver
    configuration. Returns a :class:`PreparedSolve`."""
    eps_r = _alias("eps_r", eps_r, "er", er)
    sigma = _alias("sigma", sigma, "se", se)
    mu_r = _alias("mu_r", mu_r, "mr", mr)
    receiver_component = reciever_direction if receiver_component is None else receiver_component
    if isinstance(pml_max_velocity, bool) or getattr(pml_max_velocity, "requires_grad", False):
        raise TypeError("pml_max_velocity must be a fixed finite positive scalar.")
    v_ref = float(pml_max_velocity)
    if not math.isfinite(v_ref) or v_ref <= 0.0:
        raise ValueError("pml_max_velocity must be a finite positive scalar.")
    if gradient_accumulator not in GRADIENT_ACCUMULATORS:
        raise ValueError("gradient_accumulator must be 'float32' or 'float64'.")

    device = resolve_device(device)
    if fdtd_order not in SUPPORTED_FDTD_ORDERS:
        raise ValueError("fdtd_order must be one of 2, 4, or 8.")
    if mode not in SUPPORTED_GRADIENT_MODES:
        raise ValueError("mode must be 2 or 3.")
    for name, flag in (("print_parameters", print_parameters), ("save_wavefield_history", save_wavefield_history)):
        if not isinstance(flag, bool):
            raise TypeError(f"{name} must be a bool.")
    out_dir = wavefield_directory(save_forward_wavefield_path)
    when = datetime.now() if out_dir is not None else None
    automatic = (isinstance(model_gradient_sampling_interval, str)
                 and model_gradient_sampling_interval.lower() == AUTO_SAMPLING_INTERVAL)
    if not automatic and (isinstance(model_gradient_sampling_interval, bool)
                          or not isinstance(model_gradient_sampling_interval, int)
                          or model_gradient_sampling_interval < 1):
        raise ValueError("model_gradient_sampling_interval must be a positive integer or 'auto'.")
    storage_dtype, compression, conversion, reduction = _storage_options(
        device, wavefield_storage_dtype, wavefield_compression, wavefield_conversion_backend,
        int8_reduction_backend, wavefield_compression_block_size, wavefield_compression_rate,
        save_wavefield_history, use_async_offload, out_dir is not None)
    rhs_mode = normalize_wavefield_rhs_history(wavefield_rhs_history)
    region = normalize_wavefield_history_region(wavefield_history_region)
    if source_direction not in FIELD_COMPONENTS or receiver_component not in FIELD_COMPONENTS:
        raise ValueError("source_direction and receiver_component must be 0, 1, or 2.")
    if getattr(mu_r, "requires_grad", False):
        raise NotImplementedError("DeepGPR does not currently return relative-permeability gradients.")
    lib = require_library(device)

    spacing = normalize_grid_spacing(dx)
    prepared = PreparedModel(*initialization(device, eps_r, sigma, mu_r, source_amplitudes, source_location,
                                             receiver_location, spacing, dt, pmlthick, fdtd_order))
    eps_ext, sig_ext = prepared.er, prepared.se
    nx, ny, nz, nt = prepared.nx, prepared.ny, prepared.nz, prepared.nt
    nstep, nsr, nrx = prepared.nstep, prepared.nsr, prepared.nrx
    pml_t = prepared.pmlthick
    waveforms = prepared.source_amplitudes
    src_loc = shift_locations(source_location, pml_t, device)
    rx_loc = shift_locations(receiver_location, pml_t, device)
    block, reduction = _int8_block(compression, wavefield_compression_block_size, prepared.mode, reduction)

    needs_grad = eps_ext.requires_grad or sig_ext.requires_grad
    S = _sampling(model_gradient_sampling_interval, automatic, waveforms, dt, save_wavefield_history and needs_grad)
    stores = save_wavefield_history and needs_grad
    if stores and region == "physical" and bool(pml_t.any()):
        raise ValueError("wavefield_history_region='physical' omits PML material sensitivities. "
                         "Use 'extended' for material gradients with PML.")
    lossless = compression == "none" and storage_dtype == torch.float32
    reconstruct = resolve_rhs_reconstruction(
        rhs_mode, sampling_interval=S, stores_model_history=stores,
        library_supported=library_supports(lib, "deepgpr_supports_rhs_reconstruction"), lossless_storage=lossless)
    if stores and mode == 2 and (prepared.mode != 2 or source_direction != 2 or receiver_component != 2):
        raise ValueError("mode=2 is an exact model-gradient mode only for 2D Ez-TM modeling. "
                         "Use mode=3 for 3D or other electric-field components.")

    electric = create_or_separate(E, nx, ny, nz, nstep, device, prepared.dtype)
    magnetic = create_or_separate(H, nx, ny, nz, nstep, device, prepared.dtype)
    pml_arrays = build_pml_coeffs(eps_ext, prepared.mr, dt, spacing, nx, ny, nz, pml_t, device, prepared.dtype,
                                  max_velocity=v_ref)
    descriptors, coefficients = pml_arrays[:6], pml_arrays[6:]
    phi = build_pml_phi(*descriptors, nstep, PML, device)
    tm2d = _tm2d_applies(device, prepared.mode, source_direction, (*electric, *magnetic, *phi),
                         E is not None or H is not None or PML is not None)
    config = SolverConfig(
        device=device, dtype=prepared.dtype, grid_spacing=spacing, dt=float(dt), nx=nx, ny=ny, nz=nz, nt=nt,
        nstep=nstep, nsr=nsr, nrx=nrx, pmlthick=pml_t, source_direction=source_direction,
        receiver_component=receiver_component, sampling_interval=S,
        storage=WavefieldStorageConfig(storage_dtype, compression, block, conversion, reduction),
        save_wavefield_history=save_wavefield_history, use_async_offload=bool(use_async_offload),
        fdtd_order=fdtd_order, mode=mode, debug=bool(debug), reconstruct_rhs=reconstruct, history_region=region,
        tm2d_fast_path=tm2d, grad_accumulator=gradient_accumulator)
    if print_parameters:
        from .memory import format_compute_preview
        print(format_compute_preview(config, source_waveforms=int(waveforms.shape[0]),
                                     er_requires_grad=bool(eps_ext.requires_grad),
                                     se_requires_grad=bool(sig_ext.requires_grad), save_directory=out_dir))
    tensors = SolverTensors(mu_r_pad=prepared.mr, eps_r_pad=prepared.ere, sigma_pad=prepared.see,
                            source_location=src_loc, receiver_location=rx_loc, pml_descriptors=tuple(descriptors),
                            pml_coefficients=tuple(coefficients))
    return PreparedSolve(config, tensors, eps_ext, sig_ext, waveforms, tuple(electric), tuple(magnetic), tuple(phi),
                         out_dir, when, mode, region, block, S)


def compute(device, dx=None, dt=None, source_amplitudes=None, source_location=None, receiver_location=None,
            er=None, se=None, mr=None, E=None, H=None, PML=None, pmlthick=DEFAULT_PML_THICKNESS,
            source_direction=DEFAULT_SOURCE_COMPONENT, reciever_direction=DEFAULT_RECEIVER_COMPONENT,
            model_gradient_sampling_interval=DEFAULT_GRADIENT_SAMPLING_INTERVAL,
            wavefield_storage_dtype=torch.float32, use_async_offload=False, fdtd_order=DEFAULT_FDTD_ORDER,
            mode=DEFAULT_GRADIENT_MODE, debug=False, print_parameters=False, save_forward_wavefield_path=None, *,
            eps_r=None, sigma=None, mu_r=None, receiver_component=None, save_wavefield_history=True,
            wavefield_compression=DEFAULT_WAVEFIELD_COMPRESSION, wavefield_compression_block_size=None,
            wavefield_conversion_backend=DEFAULT_CONVERSION_BACKEND,
            int8_reduction_backend=DEFAULT_INT8_REDUCTION_BACKEND, wavefield_compression_rate=None,
            wavefield_rhs_history=DEFAULT_WAVEFIELD_RHS_HISTORY,
            wavefield_history_region=DEFAULT_WAVEFIELD_HISTORY_REGION, pml_max_velocity=SPEED_OF_LIGHT,
            gradient_accumulator=DEFAULT_GRADIENT_ACCUMULATOR):
    """2D/3D FDTD forward modelling of Maxwell's equations, differentiable with
    respect to ``eps_r``, ``sigma``, ``source_amplitudes`` and the initial
    states through the exact discrete adjoint.

    Every argument has the meaning and default of the reference ``compute``.
    ``gradient_accumulator`` (keyword-only, this build) selects the precision
    of the time-summed material gradients: ``"float64"`` (default) or
    ``"float32"`` (bitwise the reference's accumulation).

    Returns ``(E_saved, (Ex, Ey, Ez), (Hx, Hy, Hz), PML, receiver_amplitudes)``.
    """
    p = prepare_compute(device, dx, dt, source_amplitudes, source_location, receiver_location, er, se, mr, E, H,
                        PML, pmlthick, source_direction, reciever_direction, model_gradient_sampling_interval,
                        wavefield_storage_dtype, use_async_offload, fdtd_order, mode, debug, print_parameters,
                        save_forward_wavefield_path, eps_r=eps_r, sigma=sigma, mu_r=mu_r,
                        receiver_component=receiver_component, save_wavefield_history=save_wavefield_history,
                        wavefield_compression=wavefield_compression,
                        wavefield_compression_block_size=wavefield_compression_block_size,
                        wavefield_conversion_backend=wavefield_conversion_backend,
                        int8_reduction_backend=int8_reduction_backend,
                        wavefield_compression_rate=wavefield_compression_rate,
                        wavefield_rhs_history=wavefield_rhs_history,
                        wavefield_history_region=wavefield_history_region, pml_max_velocity=pml_max_velocity,
                        gradient_accumulator=gradient_accumulator)
    config, tensors, out_dir, when, mode, region, block, S = (p.config, p.tensors, p.out_dir, p.when, p.mode,
                                                               p.region, p.block, p.sampling_interval)
    nx, ny, nz, nt, nstep = config.nx, config.ny, config.nz, config.nt, config.nstep
    outputs = DeepGPR.apply(config, tensors, p.eps_ext, p.sig_ext, p.waveforms, *p.electric, *p.magnetic, *p.phi)
    states, e_saved, receivers = outputs[:30], outputs[30], outputs[31]

    if out_dir is not None:
        meta = None
        origin, cells = history_box(nx, ny, nz, config.pml, region, block)
        shape = saved_history_shape(mode, saved_time_steps(nt, S), nstep, *cells)
        if save_wavefield_history and config.storage.compression == "int8":
            meta = {"compression": "int8", "block_size": block, "uncompressed_shape": shape}
        if region == "physical":
            meta = dict(meta or {})
            meta.update(history_region="physical", history_origin=origin, pmlthick=config.pml,
                        uncompressed_shape=shape)
        save_forward_wavefield(e_saved, out_dir, when, metadata=meta)  # type: ignore[arg-type]
    return e_saved, tuple(states[0:3]), tuple(states[3:6]), tuple(states[6:30]), receivers
