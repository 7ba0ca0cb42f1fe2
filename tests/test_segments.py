"""Segment planning of the checkpointed driver (no GPU needed): the longest
multiple of the sampling interval whose solve fits the memory budget."""
import pytest
import torch

from deepgpr_amd import plan_segments, segmented_compute
from deepgpr_amd.solver.memory import solve_footprint
from deepgpr_amd.solver.segments import segment_bounds
from deepgpr_amd.solver.autograd import HistoryPool, history_pool

C3 = dict(grid=(288, 288, 224), nstep=4, pml=[16] * 6)


def _need(steps, S=4):
    return solve_footprint(*C3["grid"], steps, C3["nstep"], C3["pml"], mode=3, sampling_interval=S,
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
    with pytest.raises(ValueError):
        segmented_compute("cpu", 0.02, 1e-11, z, z, z, eps_r=z, sigma=z, save_forward_wavefield_path="/tmp/x")
    with pytest.raises(ValueError, match="at least one time step"):
        segmented_compute("cpu", 0.02, 1e-11, torch.zeros(1, 0, 1), z, z, eps_r=z, sigma=z)


def test_history_pool_hands_out_aligned_views_of_one_buffer():
    pool = HistoryPool("cpu")
    a, b = pool.take([((3, 5, 7), torch.float32), ((33,), torch.int8)])
    assert a.shape == (3, 5, 7) and a.dtype == torch.float32 and a.is_contiguous()
    assert b.shape == (33,) and b.dtype == torch.int8
    buf = pool.buffer
    assert a.untyped_storage().data_ptr() == buf.untyped_storage().data_ptr() == b.untyped_storage().data_ptr()
    assert (b.data_ptr() - buf.data_ptr()) % HistoryPool.ALIGN == 0 and b.data_ptr() - a.data_ptr() >= a.nbytes
    a.fill_(1.5)
    b.fill_(-3)  # the two views do not overlap
    assert bool((a == 1.5).all()) and bool((b == -3).all())
    # a request no larger reuses the buffer; a larger one replaces it
    c, = pool.take([((4, 4), torch.bfloat16)])
    assert pool.buffer is buf and c.data_ptr() == buf.data_ptr()
    d, e = pool.take([((40, 40), torch.float32), ((40, 40), torch.float32)])
    assert pool.buffer is not buf and pool.buffer.numel() >= d.nbytes + e.nbytes
    assert e.data_ptr() - d.data_ptr() >= d.nbytes
    # keep false: the views still share one buffer, the pool no longer holds it
    pool.keep = False
    f, g = pool.take([((8,), torch.float32), ((8,), torch.float32)])
    assert pool.buffer is None and f.untyped_storage().data_ptr() == g.untyped_storage().data_ptr()


def _config(storage_dtype=torch.float32, compression="none", block=None, reconstruct=False):
    from deepgpr_amd.solver.autograd import SolverConfig
    from deepgpr_amd.solver.storage import WavefieldStorageConfig
    st = WavefieldStorageConfig(storage_dtype, compression, block, "auto", "auto")
    return SolverConfig(device=torch.device("cpu"), dtype=torch.float32, grid_spacing=(0.02, 0.02, 0.02), dt=1e-11,
                        nx=12, ny=10, nz=8, nt=9, nstep=2, nsr=1, nrx=1, pmlthick=torch.zeros(6, dtype=torch.int64),
                        source_direction=2, receiver_component=2, sampling_interval=2, storage=st,
                        save_wavefield_history=True, use_async_offload=False, fdtd_order=2, mode=3, debug=False,
                        reconstruct_rhs=reconstruct)


@pytest.mark.parametrize("fmt", ["float32 E+R", "float32 E-only", "bfloat16 E+R", "int8 E+R", "int8 E-only"])
def test_histories_from_the_pool_match_allocated_ones(fmt):
    """_allocate_histories inside a history_pool block returns tensors of the
    shapes and dtypes it allocates outside one, as views of the pool's buffer."""
    from deepgpr_amd.solver.autograd import _allocate_histories
    dtype = torch.bfloat16 if fmt.startswith("bfloat16") else torch.float32
    int8 = fmt.startswith("int8")
    cfg = _config(dtype, "int8" if int8 else "none", (4, 4, 4) if int8 else None)
    e_only = fmt.endswith("E-only")
    plain = _allocate_histories(cfg, True, e_only)
    pool = HistoryPool("cpu")
    with history_pool(pool):
        pooled = _allocate_histories(cfg, True, e_only)
    assert plain[2] == pooled[2]
    for a, b in zip(plain[:2], pooled[:2]):
        assert a.shape == b.shape and a.dtype == b.dtype and a.numel() > 0
        assert b.untyped_storage().data_ptr() == pool.buffer.untyped_storage().data_ptr()
    if e_only:  # one-frame final-E buffer
        assert pooled[1].numel() * 5 == pooled[0].numel() or int8
    # a pool made for another device is not used
    with history_pool(HistoryPool("meta")):
        other = _allocate_histories(cfg, True, e_only)
    assert other[0].device.type == "cpu"
    # without a model gradient: no R buffer, and outside the block no pool
    e, r, _ = _allocate_histories(cfg, False, False)
    assert r.numel() == 0 and e.untyped_storage().data_ptr() != pool.buffer.untyped_storage().data_ptr()


@pytest.mark.parametrize("nt,steps,S", [(2000, 388, 4), (35, 12, 1), (35, 12, 2), (36, 12, 3), (10, 12, 1),
                                        (2001, 388, 4), (41, 8, 4)])
def test_segment_bounds_cover_the_axis_on_sampling_starts(nt, steps, S):
    b = segment_bounds(nt, steps, S)
    assert b[0][0] == 0 and b[-1][1] == nt
    assert all(x[1] == y[0] for x, y in zip(b, b[1:]))  # contiguous
    assert all(a % S == 0 for a, _ in b)  # the unsegmented solve's saved steps
    assert all(0 < e - a <= steps for a, e in b)
    rem = nt % steps
    # the last segment is a full one whenever the remainder can go first
    if rem and rem % S == 0 and nt > steps:
        assert b[0][1] - b[0][0] == rem and b[-1][1] - b[-1][0] == steps
    assert segment_bounds(2000, 388, 4) == [(0, 60)] + [(60 + 388 * q, 448 + 388 * q) for q in range(5)]
