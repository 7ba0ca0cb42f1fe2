"""Segment planning of the checkpointed driver (no GPU needed): the longest
multiple of the sampling interval whose solve fits the memory budget."""
import pytest
import torch

from deepgpr_amd import estimate_compute_memory, plan_segments, segmented_compute

C3 = dict(grid=(288, 288, 224), nstep=4, pml=[16] * 6)


def _need(steps, S=4):
    return estimate_compute_memory(*C3["grid"], steps, C3["nstep"], C3["pml"], mode=3, sampling_interval=S,
                                   nrx=256)["total"]


def test_c3_four_shots_fp32_history_needs_segments():
    assert _need(2000) > 288e9  # four shots' FP32 E+R history exceeds one MI355X
    budget = 250 * 10**9
    steps = plan_segments(2000, budget, C3["grid"], C3["nstep"], C3["pml"], mode=3, sampling_interval=4, nrx=256)
    assert steps % 4 == 0 and 4 <= steps < 2000
    assert _need(steps) <= budget < _need(steps + 4)


def test_plan_takes_the_whole_axis_when_it_fits_and_rejects_a_tiny_budget():
    assert plan_segments(2000, 10**13, C3["grid"], C3["nstep"], C3["pml"], sampling_interval=4) == 2000
    assert plan_segments(37, 10**13, C3["grid"], C3["nstep"], C3["pml"], sampling_interval=4) == 37
    with pytest.raises(ValueError):
        plan_segments(2000, 10**6, C3["grid"], C3["nstep"], C3["pml"], sampling_interval=4)


def test_segmented_compute_refuses_initial_states():
    z = torch.zeros(1)
    with pytest.raises(ValueError):
        segmented_compute("cpu", 0.02, 1e-11, z, z, z, eps_r=z, sigma=z, E=(z, z, z))
