"""torch.autograd bridge to the HIP forward solver and its exact discrete adjoint
(src/DeepGPR/solver/autograd.py:63-760).

The 30 state tensors (E, H and 24 CPML auxiliaries) are inputs and outputs at
the same time: the kernels advance them in place, which is what lets callers
chain segments with torch.utils.checkpoint. Backward clones the incoming state
cotangents (the adjoint advances them in place) so caller-owned gradients are
never modified.
"""
from __future__ import annotations

import contextlib
import contextvars
from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence, Tuple

import torch

from ..config.constants import CPML_FACE_NAMES, TM2D_INACTIVE_PHI_INDICES
from ..native.bridge import invoke
from ..native.loader import library_supports, require_library
from ..utils.devices import resolve_device
from ..utils.exceptions import NativeLibraryError, NonFiniteTensorError
from .storage import (
    WavefieldStorageConfig, decompress_wavefield_history, encode_storage_type, history_spatial_shape,
    int8_history_layout, saved_history_shape, saved_time_steps,
)

STATE_TENSOR_COUNT = 30
STATE_NAMES: Tuple[str, ...] = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz") + tuple(
    f"{face}{field}Phi{k}" for face in CPML_FACE_NAMES for field in ("E", "H") for k in (1, 2))
TM2D_INACTIVE_STATE_INDICES: Tuple[int, ...] = (0, 1, 5) + tuple(6 + k for k in TM2D_INACTIVE_PHI_INDICES)


@dataclass(frozen=True)
class SolverConfig:
    """Scalar configuration of one solve (see the reference SolverConfig)."""

    device: torch.device
    dtype: torch.dtype
    grid_spacing: Tuple[float, float, float]
    dt: float
    nx: int
    ny: int
    nz: int
    nt: int
    nstep: int
    nsr: int
    nrx: int
    pmlthick: torch.Tensor
    source_direction: int
    receiver_component: int
    sampling_interval: int
    storage: WavefieldStorageConfig
    save_wavefield_history: bool
    use_async_offload: bool
    fdtd_order: int
    mode: int
    debug: bool
    reconstruct_rhs: bool = False
    history_region: str = "extended"
    tm2d_fast_path: bool = False
    grad_accumulator: str = "float64"

    @property
    def pml(self) -> List[int]:
        return [int(self.pmlthick[i]) for i in range(6)]

    @property
    def history_cells(self) -> Tuple[int, int, int]:
        block = self.storage.block_size if self.storage.compression == "int8" else None
        return history_spatial_shape(self.nx, self.ny, self.nz, self.pml, self.history_region, block)


@dataclass(frozen=True)
class SolverTensors:
    mu_r_pad: torch.Tensor
    eps_r_pad: torch.Tensor
    sigma_pad: torch.Tensor
    source_location: torch.Tensor
    receiver_location: torch.Tensor
    pml_descriptors: Tuple[torch.Tensor, ...]
    pml_coefficients: Tuple[torch.Tensor, ...]  # x0_e, x0_h, xm_e, xm_h, ...


def _native_pml_coefficients(coefficients: Tuple[torch.Tensor, ...]) -> List[torch.Tensor]:
    """Face-interleaved (e, h) pairs -> native order (6 electric, 6 magnetic)."""
    return [coefficients[2 * f] for f in range(6)] + [coefficients[2 * f + 1] for f in range(6)]


def _common(config: SolverConfig) -> Dict[str, Any]:
    dx, dy, dz = config.grid_spacing
    return {
        "fdtd_order": config.fdtd_order, "pml": config.pml, "dt": config.dt, "nt": config.nt,
        "nshot": config.nstep, "nreceiver": config.nrx, "dx": dx, "dy": dy, "dz": dz,
        "nx_fields": config.nx + 1, "ny_fields": config.ny + 1, "nz_fields": config.nz + 1,
        "receiver_component": config.receiver_component, "nsource": config.nsr,
        "source_component": config.source_direction, "sampling_interval": config.sampling_interval,
        "fwi_mode": config.mode, "accumulate_fp64": int(config.grad_accumulator == "float64"),
    }


def _update_coefficients(config: SolverConfig) -> List[torch.Tensor]:
    buf = torch.zeros((6, config.nx + 1, config.ny + 1, config.nz + 1), device=config.device, dtype=config.dtype)
    return list(buf.unbind(0))


def _check_capabilities(lib: Any, config: SolverConfig) -> None:
    needs = []
    if config.tm2d_fast_path:
        needs.append("deepgpr_supports_tm2d_fast_path")
    if config.history_region == "physical":
        needs.append("deepgpr_supports_physical_history")
    if config.reconstruct_rhs:
        needs.append("deepgpr_supports_rhs_reconstruction")
    if config.save_wavefield_history and config.storage.compression == "int8":
        needs.append("deepgpr_supports_int8_wavefield")
    for cap in needs:
        if not library_supports(lib, cap):
            raise NativeLibraryError(f"The loaded HIP library lacks {cap}; rebuild it.")


class HistoryPool:
    """One device buffer that solves run one after another take their
    histories from, instead of allocating them each time.

    Successive histories of the same size then reuse the same bytes: no
    release and re-allocation of blocks of up to hundreds of GB between the
    solves (which costs seconds each on this device) and no caching-allocator
    blocks split by smaller allocations in between. Each solve's histories
    are views of the same bytes, so a solve may take them only once the
    previous solve's backward has been queued on the same stream; the
    segment driver, which re-runs one segment at a time inside its
    checkpointed backward, is the user it is made for. The buffer grows to
    the largest request and the pool holds it while `keep` is true; with
    `keep` false the next take() hands out its views and lets go of it.

    On a GPU the buffer lives in a private pool of torch's caching allocator
    (torch.cuda.MemPool), so no other allocation is ever carved from it:
    once the pool has let go of the buffer and the views are gone, its
    memory is released whole at the allocator's next empty_cache() or
    out-of-memory retry instead of staying pinned by small tensors placed
    in the freed block. A new buffer that does not fit the device's free
    memory empties the allocator's cache first."""

    ALIGN = 256

    def __init__(self, device):
        self.device = resolve_device(device)
        self.buffer: Optional[torch.Tensor] = None
        self.keep = True
        self._mem = None  # the private allocator pool of the current buffer

    def _allocate(self, nbytes: int) -> torch.Tensor:
        if self.device.type != "cuda":
            return torch.empty(nbytes, device=self.device, dtype=torch.uint8)
        if torch.cuda.mem_get_info(self.device)[0] < nbytes:
            # an allocation routed to a private pool does not make the
            # allocator release its cache when it fails, so release it first
            torch.cuda.empty_cache()
        self._mem = torch.cuda.MemPool()  # a new pool per buffer: the old one is released with its buffer
        with torch.cuda.use_mem_pool(self._mem, self.device):
            return torch.empty(nbytes, device=self.device, dtype=torch.uint8)

    def take(self, specs: Sequence[Tuple[Tuple[int, ...], torch.dtype]]) -> List[torch.Tensor]:
        sizes = []
        for shape, dtype in specs:
            n = 1
            for d in shape:
                n *= int(d)
            sizes.append(n * torch.empty(0, dtype=dtype).element_size())
        offsets, total = [], 0
        for n in sizes:
            offsets.append(total)
            total += -(-n // self.ALIGN) * self.ALIGN
        if self.buffer is None or self.buffer.numel() < total:
            self.buffer = self._mem = None  # release the smaller buffer before the larger one is allocated
            self.buffer = self._allocate(max(total, 1))
        views = [self.buffer[o:o + n].view(dtype).view(shape)
                 for o, n, (shape, dtype) in zip(offsets, sizes, specs)]
        if not self.keep:
            self.buffer = self._mem = None
        return views


_HISTORY_POOL: contextvars.ContextVar[Optional[HistoryPool]] = contextvars.ContextVar("deepgpr_amd_history_pool",
                                                                                     default=None)


@contextlib.contextmanager
def history_pool(pool: Optional[HistoryPool]):
    """Solves started inside the block take their histories from `pool`."""
    token = _HISTORY_POOL.set(pool)
    try:
        yield pool
    finally:
        _HISTORY_POOL.reset(token)


def _allocate_histories(config: SolverConfig, needs_model_gradient: bool, reconstruct_rhs: bool):
    """E_saved and R_saved (or the one-frame final-E buffer of an E-only history).

    Histories always stay resident in device memory (use_async_offload is
    accepted for API compatibility; 288 GB of HBM holds what the reference
    staged to pinned host memory). Inside a history_pool block they are views
    of the pool's buffer."""
    dev, st = config.device, config.storage
    nt_saved = saved_time_steps(config.nt, config.sampling_interval)
    shape = saved_history_shape(config.mode, nt_saved, config.nstep, *config.history_cells)
    if not config.save_wavefield_history:
        return torch.empty(0, device=dev, dtype=st.dtype), torch.empty(0, device=dev, dtype=st.dtype), shape
    if st.compression == "int8":
        n = int8_history_layout(shape, st.block_size)["packed_bytes"]  # type: ignore[arg-type]
        frame: Tuple[Tuple[int, ...], torch.dtype] = ((n,), torch.int8)
    else:
        frame = (tuple(shape), st.dtype)
    final = reconstruct_rhs and needs_model_gradient
    specs = [frame]
    if final:
        fshape = saved_history_shape(config.mode, 1, config.nstep, *config.history_cells)
        if st.compression == "int8":
            specs.append(((int8_history_layout(fshape, st.block_size)["packed_bytes"],), torch.int8))
        else:
            specs.append((tuple(fshape), st.dtype))
    elif needs_model_gradient:
        specs.append(frame)
    pool = _HISTORY_POOL.get()
    if pool is not None and pool.device == resolve_device(dev):
        got = pool.take(specs)
    else:
        got = [torch.empty(sh, device=dev, dtype=dt) for sh, dt in specs]
    if len(got) == 1:
        got.append(torch.empty(0, device=dev, dtype=frame[1]))
    return got[0], got[1], shape


def check_tensors_for_nan_inf(d: str = "", **tensors: torch.Tensor) -> None:
    """Raise :class:`NonFiniteTensorError` naming every tensor with NaN/Inf."""
    bad = [name for name, t in tensors.items()
           if isinstance(t, torch.Tensor) and t.numel() and t.is_floating_point() and not bool(torch.isfinite(t).all())]
    if bad:
        raise NonFiniteTensorError(f"Non-finite values after {d}: {', '.join(bad)}")


def _own(gradient: Optional[torch.Tensor], shape, device, dtype) -> torch.Tensor:
    if gradient is None:
        return torch.zeros(shape, device=device, dtype=dtype)
    return gradient.clone(memory_format=torch.contiguous_format)


def forward_native_args(config: SolverConfig, tensors: SolverTensors, source_amplitudes: torch.Tensor,
                        states: Sequence[torch.Tensor], needs_grad: bool):
    """The dgx_forward arguments of one solve, and the arrays it fills
    (histories, receivers). Shared by the autograd function and the batched
    solver."""
    dev = config.device
    src_loc = tensors.source_location.to(device=dev, dtype=torch.int32).contiguous()
    rx_loc = tensors.receiver_location.to(device=dev, dtype=torch.int32).contiguous()
    if (config.save_wavefield_history and needs_grad and config.history_region == "physical"
            and bool(config.pmlthick.any())):
        raise ValueError("wavefield_history_region='physical' omits PML material sensitivities. "
                         "Use 'extended' for material gradients with PML.")
    reconstruct = bool(config.reconstruct_rhs and config.save_wavefield_history and needs_grad)
    storage_type = encode_storage_type(config.storage, config.save_wavefield_history, rhs_from_e=reconstruct,
                                       physical_history=config.history_region == "physical",
                                       tm2d=config.tm2d_fast_path)
    e_saved, r_saved, history_shape = _allocate_histories(config, needs_grad, reconstruct)
    receivers = torch.empty((config.nstep, config.nt, config.nrx), device=dev, dtype=config.dtype)
    args = _common(config)
    args.update(
        eps_r_pad=tensors.eps_r_pad, sigma_pad=tensors.sigma_pad, mu_r_pad=tensors.mu_r_pad,
        E_saved=e_saved, R_saved=r_saved, E=list(states[0:3]), H=list(states[3:6]),
        update_coeff=_update_coefficients(config), phi=list(states[6:30]),
        pml_coeff=_native_pml_coefficients(tensors.pml_coefficients),
        receiver_location=rx_loc, receiver_data=receivers, source_location=src_loc,
        source_waveform=source_amplitudes, storage_type=storage_type,
        save_model_history=int(config.save_wavefield_history and needs_grad),
        use_async_offload=int(config.use_async_offload),
    )
    out = {"E_saved": e_saved, "R_saved": r_saved, "history_shape": history_shape, "receivers": receivers,
           "reconstruct": reconstruct}
    return args, out


class DeepGPR(torch.autograd.Function):
    """``DeepGPR.apply(config, tensors, eps_r, sigma, source_amplitudes, *states)``
    returns the 30 advanced states, ``E_saved`` and the receiver data
    ``(nstep, nt, nrx)``."""

    @staticmethod
    def forward(ctx: Any, config: SolverConfig, tensors: SolverTensors, eps_r: torch.Tensor,  # type: ignore[override]
                sigma: torch.Tensor, source_amplitudes: torch.Tensor, *states: torch.Tensor):
        if len(states) != STATE_TENSOR_COUNT:
            raise ValueError(f"Expected {STATE_TENSOR_COUNT} state tensors, got {len(states)}.")
        ctx.set_materialize_grads(False)
        dev = config.device
        lib = require_library(dev)
        _check_capabilities(lib, config)
        source_amplitudes = source_amplitudes.contiguous()
        want_eps, want_sig = bool(eps_r.requires_grad), bool(sigma.requires_grad)
        needs_grad = want_eps or want_sig
        args, out = forward_native_args(config, tensors, source_amplitudes, states, needs_grad)
        src_loc, rx_loc = args["source_location"], args["receiver_location"]
        storage_type, reconstruct = args["storage_type"], out["reconstruct"]
        e_saved, r_saved, history_shape, receivers = out["E_saved"], out["R_saved"], out["history_shape"], \
            out["receivers"]
        if config.save_wavefield_history:
            ctx.save_for_backward(tensors.mu_r_pad, src_loc, rx_loc, *tensors.pml_coefficients,
                                  tensors.eps_r_pad, tensors.sigma_pad)
        else:
            ctx.save_for_backward()
        ctx.config = config
        ctx.want_eps, ctx.want_sig = want_eps, want_sig
        ctx.want_src = bool(source_amplitudes.requires_grad)
        ctx.source_shape = tuple(source_amplitudes.shape)
        ctx.storage_type = storage_type
        ctx.reconstruct = reconstruct
        ctx.state_shapes = tuple(s.shape for s in states)
        ctx.receiver_shape = (config.nstep, config.nt, config.nrx)

        invoke("forward", args, dev)
        if config.debug:
            int8 = config.storage.compression == "int8"
            block = config.storage.block_size or (1, 1, 1)
            de = decompress_wavefield_history(e_saved, history_shape, block) if int8 and e_saved.numel() else e_saved
            check_tensors_for_nan_inf(d="forward", E_saved=de, **dict(zip(STATE_NAMES, states)))
        ctx.E_saved = e_saved
        ctx.R_saved = None if reconstruct else r_saved
        ctx.E_final = r_saved if reconstruct else None
        ctx.mark_non_differentiable(e_saved)
        return (*states, e_saved, receivers)

    @staticmethod
    def backward(ctx: Any, *grad_outputs: Optional[torch.Tensor]):  # type: ignore[override]
        config: SolverConfig = ctx.config
        if not config.save_wavefield_history:
            raise RuntimeError("Forward wavefield history is disabled; adjoint/model-gradient backward is unavailable.")
        if ctx.E_saved is None:
            raise RuntimeError("DeepGPR's saved forward wavefield was already released by a previous backward pass. "
                               "Re-run compute before calling backward again.")
        dev, dtype = config.device, config.dtype
        data_grad = grad_outputs[STATE_TENSOR_COUNT + 1]
        saved = tuple(t.contiguous() for t in ctx.saved_tensors)
        mu_r_pad, src_loc, rx_loc = saved[:3]
        coefficients = saved[3:15]
        eps_r_pad, sigma_pad = saved[15:17]
        e_saved = ctx.E_saved.contiguous()
        r_saved = (ctx.E_final if ctx.reconstruct else ctx.R_saved).contiguous()
        lambdas = [_own(g, shape, dev, dtype) for g, shape in zip(grad_outputs[:STATE_TENSOR_COUNT], ctx.state_shapes)]
        data_grad = torch.zeros(ctx.receiver_shape, device=dev, dtype=dtype) if data_grad is None else data_grad.contiguous()
        cells = (config.nx, config.ny, config.nz)
        g_eps = torch.zeros(cells, device=dev, dtype=dtype) if ctx.want_eps else torch.empty(0, device=dev, dtype=dtype)
        g_sig = torch.zeros(cells, device=dev, dtype=dtype) if ctx.want_sig else torch.empty(0, device=dev, dtype=dtype)
        g_src = (torch.zeros(ctx.source_shape, device=dev, dtype=dtype) if ctx.want_src
                 else torch.empty(0, device=dev, dtype=dtype))
        args = _common(config)
        args.update(
            eps_r_pad=eps_r_pad, sigma_pad=sigma_pad, mu_r_pad=mu_r_pad, E_saved=e_saved, R_saved=r_saved,
            E=lambdas[0:3], H=lambdas[3:6], update_coeff=_update_coefficients(config), phi=lambdas[6:30],
            pml_coeff=_native_pml_coefficients(coefficients), receiver_location=rx_loc, data_grad=data_grad,
            source_location=src_loc, grad_source=g_src, source_requires_grad=int(ctx.want_src),
            grad_eps_r=g_eps, grad_sigma=g_sig, eps_r_requires_grad=int(ctx.want_eps),
            sigma_requires_grad=int(ctx.want_sig), storage_type=ctx.storage_type, use_async_offload=0,
        )
        invoke("backward", args, dev)
        if config.tm2d_fast_path:
            # Ex, Ey, Hz and their CPML auxiliaries never couple back into the
            # 2D TM fields: their initial-state cotangents are zero.
            for k in TM2D_INACTIVE_STATE_INDICES:
                lambdas[k].zero_()
        need = ctx.needs_input_grad[5:5 + STATE_TENSOR_COUNT]
        if config.debug:
            check = {STATE_NAMES[k]: lambdas[k] for k in range(STATE_TENSOR_COUNT) if need[k]}
            if ctx.want_eps:
                check["grad_eps_r"] = g_eps
            if ctx.want_sig:
                check["grad_sigma"] = g_sig
            if ctx.want_src:
                check["grad_source"] = g_src
            check_tensors_for_nan_inf(d="backward", **check)
        ctx.E_saved = ctx.R_saved = ctx.E_final = None
        return (None, None, g_eps if ctx.want_eps else None, g_sig if ctx.want_sig else None,
                g_src if ctx.want_src else None, *(g if n else None for g, n in zip(lambdas, need)))
