"""Segment-checkpoint driver: compute() over consecutive segments of the time
axis, each under torch.utils.checkpoint, so that only one segment's wavefield
history is alive at a time (the pattern of the reference's
tests/checkpoint_example.ipynb cell 5 and tests/test_external_pml.py:145-191,
as a library function).

The forward runs the segments in order without autograd, keeping only their
boundary states (E, H and the 24 CPML auxiliaries), except the last, which
it runs with its history. The backward of each segment re-runs its forward
with the history, then its exact adjoint, so at most one segment's history
is in memory at a time; the histories all live in one HistoryPool buffer. Receivers and final states
equal those of one unsegmented compute() bit for bit; the eps_r, sigma and
source gradients equal its gradients up to float32 summation order (autograd
adds per-segment sums where the unsegmented adjoint sums all steps in one
accumulator). A segment length is a multiple of the gradient sampling
interval, so the saved steps are those of the unsegmented solve.
"""
from __future__ import annotations

import math
from typing import Optional

import torch
from torch.utils.checkpoint import checkpoint

from ..config.defaults import (DEFAULT_FDTD_ORDER, DEFAULT_GRADIENT_ACCUMULATOR, DEFAULT_GRADIENT_MODE,
                               DEFAULT_GRADIENT_SAMPLING_INTERVAL, DEFAULT_PML_THICKNESS)
from ..preprocessing.model_setup import initialization
from .autograd import HistoryPool, history_pool
from .fields import checkpoint_initial_field
from .memory import solve_footprint
from .modeling import compute
from ..utils.devices import resolve_device

_STORAGE_NAMES = {torch.float32: "float32", torch.float16: "float16", torch.bfloat16: "bfloat16"}


def plan_segments(nt: int, budget_bytes: int, grid, nstep: int, pml, *, mode: int = 3, sampling_interval: int = 1,
                  storage: str = "float32", compression: str = "none", block_size=(4, 4, 4),
                  reconstruct_rhs: bool = False, fp64_accumulators: bool = True, nrx: int = 1) -> int:
    """Steps per segment: the longest multiple of `sampling_interval` whose
    solve (solve_footprint over the extended grid `grid` = (nx, ny,
    nz)) fits `budget_bytes`, at most nt. Raises ValueError when not even one
    sampling interval fits."""
    S = max(int(sampling_interval), 1)

    def need(steps: int) -> int:
        return solve_footprint(*grid, steps, nstep, pml, mode=mode, sampling_interval=S, storage=storage,
                               compression=compression, block_size=block_size, reconstruct_rhs=reconstruct_rhs,
                               fp64_accumulators=fp64_accumulators, nrx=nrx)["total"]

    if need(S) > budget_bytes:
        raise ValueError(f"segmented_compute: one segment of {S} steps needs {need(S)} bytes, "
                         f"more than the budget of {budget_bytes}")
    lo, hi = 1, max(1, math.ceil(nt / S))  # in sampling intervals
    while lo < hi:  # the memory need grows with the segment length
        mid = (lo + hi + 1) // 2
        if need(mid * S) <= budget_bytes:
            lo = mid
        else:
            hi = mid - 1
    return min(nt, lo * S)


def segment_bounds(nt: int, segment_steps: int, sampling_interval: int = 1):
    """[start, end) of each segment of nt steps. Every segment starts on a
    multiple of the sampling interval, so the saved steps are those of the
    unsegmented solve. The remainder segment goes first when that keeps the
    rule, so that the last segment, where the backward starts and which the
    forward runs with its history instead of the backward re-running it, is
    a full one."""
    S = max(int(sampling_interval), 1)
    rem = nt % segment_steps
    if rem and rem % S == 0:
        return [(0, rem)] + [(a, a + segment_steps) for a in range(rem, nt, segment_steps)]
    return [(a, min(a + segment_steps, nt)) for a in range(0, nt, segment_steps)]


def segmented_compute(device, dx=None, dt=None, source_amplitudes=None, source_location=None,
                      receiver_location=None, *, eps_r=None, sigma=None, mu_r=None, segment_steps: Optional[int] = None,
                      memory_budget: Optional[int] = None, pmlthick=DEFAULT_PML_THICKNESS,
                      fdtd_order: int = DEFAULT_FDTD_ORDER,
                      model_gradient_sampling_interval=DEFAULT_GRADIENT_SAMPLING_INTERVAL, **compute_kwargs):
    """compute() of the whole time axis as checkpointed segments.

    segment_steps: steps per segment (rounded up to a multiple of the gradient
    sampling interval); by default the longest segment whose solve fits
    `memory_budget` bytes (default: 60 % of the device's free memory,
    counting what torch's allocator holds unused as free), from
    solve_footprint. Other keywords go to compute() unchanged (E, H
    and PML initial states are not accepted: the run starts from zero states,
    as checkpoint_initial_field gives them; nor is save_forward_wavefield_path).

    Returns (None, E, H, PML, receivers) like compute(): there is no single
    history to return.
    """
    for key in ("E", "H", "PML"):
        if compute_kwargs.get(key) is not None:
            raise ValueError(f"segmented_compute starts from zero states; {key} is not accepted")
    if compute_kwargs.get("save_forward_wavefield_path") is not None:
        raise ValueError("segmented_compute keeps no single history to save; save_forward_wavefield_path "
                         "is not accepted")
    if model_gradient_sampling_interval == "auto":
        raise ValueError("segmented_compute needs an explicit model_gradient_sampling_interval")
    device = resolve_device(device)
    nt = int(source_amplitudes.shape[1])
    if nt < 1:
        raise ValueError("segmented_compute needs at least one time step")
    S = int(model_gradient_sampling_interval)
    if segment_steps is None:
        prep = initialization(device, eps_r, sigma, mu_r, source_amplitudes, source_location, receiver_location, dx,
                              dt, pmlthick, fdtd_order)
        if memory_budget is None:
            # memory the caching allocator holds unused is released to an
            # allocation that needs it, so it counts as free; 60 % of that
            # leaves room for what the backward keeps beside one segment's
            # solve: every segment's boundary states, the gradients
            free = torch.cuda.mem_get_info(device)[0]
            cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
            memory_budget = int(0.6 * (free + cached))
        dtype = compute_kwargs.get("wavefield_storage_dtype", torch.float32)
        compression = compute_kwargs.get("wavefield_compression", "none") or "none"
        rhs = compute_kwargs.get("wavefield_rhs_history", "auto")
        e_only = rhs == "reconstructed" or (rhs == "auto" and dtype == torch.float32 and compression == "none"
                                            and S == 1)
        segment_steps = plan_segments(
            nt, memory_budget, (prep.nx, prep.ny, prep.nz), prep.nstep, [int(p) for p in prep.pmlthick],
            mode=compute_kwargs.get("mode", DEFAULT_GRADIENT_MODE), sampling_interval=S,
            storage=_STORAGE_NAMES.get(dtype, "float32"),
            compression=compression, reconstruct_rhs=e_only,
            fp64_accumulators=compute_kwargs.get("gradient_accumulator", DEFAULT_GRADIENT_ACCUMULATOR) == "float64",
            nrx=int(receiver_location.shape[1]))
    segment_steps = max(S, -(-int(segment_steps) // S) * S)
    E, H, P = checkpoint_initial_field(device, None, dx, dt, source_amplitudes, source_location, receiver_location,
                                       er=eps_r, se=sigma, mr=mu_r, pmlthick=pmlthick, fdtd_order=fdtd_order)
    state = (*E, *H, *P)
    kwargs = dict(compute_kwargs, pmlthick=pmlthick, fdtd_order=fdtd_order,
                  model_gradient_sampling_interval=model_gradient_sampling_interval)

    # the last segment's run and every re-run take their history from one
    # buffer, so the solve allocates a history once; the backward re-runs
    # the first segment last, and that re-run lets the pool go of the
    # buffer, which is freed when that segment's backward is done
    pool = HistoryPool(device)

    # only runs with grad enabled (the last segment's, the backward's
    # re-runs) store a history; the checkpointed forward runs store none
    # (histories stay in HBM, so offload is moot)
    forward_pass = dict(kwargs, save_wavefield_history=False, use_async_offload=False)

    def segment(first, er, se, wave, *boundary):
        pool.keep = not first
        work = tuple(t.clone() for t in boundary)
        with history_pool(pool):
            r = compute(device, dx, dt, wave, source_location, receiver_location, eps_r=er, sigma=se, mu_r=mu_r,
                        E=work[:3], H=work[3:6], PML=work[6:],
                        **(kwargs if torch.is_grad_enabled() else forward_pass))
        return (*r[1], *r[2], *r[3], r[-1])

    bounds = segment_bounds(nt, segment_steps, S)
    wants_grad = torch.is_grad_enabled() and any(
        getattr(t, "requires_grad", False) for t in (eps_r, sigma, source_amplitudes))
    # Re-entrant checkpoints run each other segment's forward without
    # autograd, so nothing keeps its history (the autograd function holds it
    # as context, which the non-re-entrant form's saved-tensor hooks do not
    # release); the backward re-runs one segment at a time with its history.
    traces = []
    for q, (a, b) in enumerate(bounds):
        wave = source_amplitudes[:, a:b]
        if q == len(bounds) - 1 and wants_grad:
            out = segment(q == 0, eps_r, sigma, wave, *state)
        else:
            out = checkpoint(segment, q == 0, eps_r, sigma, wave, *state, use_reentrant=True)
        state, data = out[:-1], out[-1]
        traces.append(data)
    return None, tuple(state[0:3]), tuple(state[3:6]), tuple(state[6:30]), torch.cat(traces, dim=1)
