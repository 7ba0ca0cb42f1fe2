"""dgx_scratch_bytes (host logic only, no GPU): the device memory a solve
needs beyond the caller's arrays, and the matching term of
solve_footprint used for admission."""
import ctypes
import os

import pytest
import torch

from deepgpr_amd.native.bridge import build_solve
from deepgpr_amd.native.loader import get_library
from deepgpr_amd.solver.memory import solve_footprint

_T = torch.zeros(8)  # stands in for every array: only the sizes are read


def _scratch(nx, ny, nz, nshot, tm=False, fp64=0):
    args = dict(nx_fields=nx, ny_fields=ny, nz_fields=nz, nshot=nshot, nt=10, fdtd_order=4, pml=[0] * 6,
                dx=0.02, dy=0.02, dz=0.02, dt=1e-11, storage_type=(1 << 6) if tm else 0, source_component=2,
                receiver_component=2, sampling_interval=1, fwi_mode=3, nsource=1, nreceiver=1, save_model_history=1,
                accumulate_fp64=fp64, E=[_T] * 3, H=[_T] * 3, eps_r_pad=_T, sigma_pad=_T, mu_r_pad=_T,
                update_coeff=[_T] * 6, source_location=_T, source_waveform=_T, receiver_location=_T, receiver_data=_T,
                data_grad=_T, grad_eps_r=_T, grad_sigma=_T, grad_source=_T)
    lib = get_library()
    solve = build_solve(args)
    back = lib.dgx_scratch_bytes(ctypes.byref(solve), 1)
    assert back >= 0, lib.dgx_last_error()
    return back


SPLIT = os.environ.get("DEEPGPR_AMD_3D_ADJOINT") == "split"
CELL = os.environ.get("DEEPGPR_AMD_3D_SWEEP") != "xmarch"


def test_3d_adjoint_scratch():
    cells = 30 * 31 * 32
    # the one-thread-per-cell adjoint keeps one or two sets of six FP32 channel
    # weights per cell and shot; the x-marching sweeps form them on the fly
    per = (24 if SPLIT else 48) if CELL else 0
    assert _scratch(30, 31, 32, 2) == per * 2 * cells
    assert _scratch(30, 31, 32, 2, fp64=1) == per * 2 * cells + 16 * 29 * 30 * 31  # plus FP64 accumulators


def test_2d_tm_adjoint_needs_no_weight_arrays():
    assert _scratch(30, 31, 2, 2, tm=True) == 0


@pytest.mark.skipif(SPLIT, reason="the estimate prices the default (fused) adjoint")
def test_estimate_matches_the_library():
    est = solve_footprint(29, 30, 31, 10, 2, [0] * 6, mode=3)
    assert est.get("adjoint_weights", 0) == _scratch(30, 31, 32, 2)
