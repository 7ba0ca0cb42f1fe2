"""The run-time tile tuner of the fused 2D Ez-TM kernels and the graphs of the
batch runtime change how the work is launched, never its results: a solve
gives the same bits with the tuned tile height as with each forced height,
and a batch replaying 32-step graphs gives the bits of the same batch
launched step by step and of one solve at a time."""
import os
import sys

import pytest
import torch

sys.path.insert(0, os.path.dirname(__file__))
from gpu_helpers import capture  # noqa: E402

pytestmark = pytest.mark.gpu


def _solve(dev):
    import deepgpr_amd as dg
    gen = torch.Generator().manual_seed(77)
    nx, ny, nt, nshot = 150, 118, 60, 3
    eps = (1.0 + 8.0 * torch.rand(nx, ny, generator=gen)).to(dev).requires_grad_(True)
    sig = (0.01 * torch.rand(nx, ny, generator=gen)).to(dev).requires_grad_(True)
    t = torch.arange(nt, dtype=torch.float32) * 3e-11
    f0 = 4e8
    a2 = (torch.pi * f0 * (t - 1.2 / f0)) ** 2
    wave = ((1 - 2 * a2) * torch.exp(-a2)).reshape(1, nt, 1).to(dev)
    src = torch.tensor([[[20, 10 + 30 * s, 0]] for s in range(nshot)], device=dev)
    rx = torch.tensor([[[25, 5 + 9 * r, 0] for r in range(10)] for _ in range(nshot)], device=dev)
    out = dg.compute(dev, dx=0.02, dt=3e-11, source_amplitudes=wave, source_location=src, receiver_location=rx,
                     eps_r=eps, sigma=sig, fdtd_order=4, gradient_accumulator="float32")
    (out[-1] ** 2).sum().backward()
    return out[-1].detach().clone(), eps.grad.clone(), sig.grad.clone()


def test_tuned_tile_rows_give_the_same_bits(cuda_device, monkeypatch):
    from deepgpr_amd.native.loader import get_library
    print("[gpu tuning] tile tuner", flush=True)
    monkeypatch.delenv("DEEPGPR_AMD_TM2D_TILE_ROWS", raising=False)
    monkeypatch.delenv("DEEPGPR_AMD_TM2D_TUNE", raising=False)
    calls = capture(monkeypatch)
    tuned = _solve(cuda_device)
    fin = calls[0][1]
    assert fin["storage_type"] & (1 << 6)  # the fused 2D Ez-TM path
    rows = get_library().dgx_tm2d_tile_rows_tuned(fin["nx_fields"], fin["ny_fields"], fin["nshot"],
                                                  fin["fdtd_order"], 1)
    assert rows in (4, 8, 16), rows
    for forced in ("4", "8", "16"):
        monkeypatch.setenv("DEEPGPR_AMD_TM2D_TILE_ROWS", forced)
        got = _solve(cuda_device)
        for a, b in zip(got, tuned):
            assert torch.equal(a, b), forced
    monkeypatch.delenv("DEEPGPR_AMD_TM2D_TILE_ROWS")
    monkeypatch.setenv("DEEPGPR_AMD_TM2D_TUNE", "0")
    for a, b in zip(_solve(cuda_device), tuned):
        assert torch.equal(a, b)


def test_batch_graphs_give_the_same_bits(cuda_device, monkeypatch):
    from deepgpr_amd import batch, serving, synthetic
    print("[gpu tuning] batch graphs", flush=True)
    reqs = synthetic.batch_instances(8, seed=21, nt_scale=0.12)
    for k, r in enumerate(reqs):
        r.rid = k
    alone = serving.run_sequential(reqs, cuda_device)["results"]

    def batched(graphs):
        monkeypatch.setenv("DEEPGPR_AMD_BATCH_GRAPHS", "1" if graphs else "0")
        with batch.BatchSolver(cuda_device) as solver:
            ids = {solver.add(r): r.rid for r in reqs[:5]}
            solver.step(64)  # whole graphs while no instance changes phase
            ids.update({solver.add(r): r.rid for r in reqs[5:]})  # joins rebuild the table
            solver.run(64)
            return {ids[n]: solver.result(n) for n in ids}

    with_graphs, without = batched(True), batched(False)
    for r in reqs:
        a = alone[r.rid]
        for got in (with_graphs[r.rid], without[r.rid]):
            rx, g_eps, g_sig = got
            assert torch.equal(rx, a.receivers), r.rid
            assert torch.equal(g_eps, a.grad_eps_r), r.rid
            assert torch.equal(g_sig, a.grad_sigma), r.rid
