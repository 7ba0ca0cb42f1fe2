"""Admission, slot and ordering behaviour of the request loop (CPU, fake solver),
and GPU checks that a result is the same alone and in the mix."""
import threading
import time

import pytest
import torch

from deepgpr_amd import serving, synthetic


def _req(rid, side=128, nt=100, gradient=False):
    eps = torch.full((side, side), 4.0)
    sig = torch.zeros((side, side))
    wave = torch.zeros((1, nt, 1))
    src = torch.tensor([[[5, side // 2, 0]]])
    rx = torch.tensor([[[5, 3, 0], [5, side - 3, 0]]])
    return serving.SolveRequest(eps, sig, wave, src, rx, 0.02, 3e-11, gradient=gradient, rid=rid)


class _Fake:
    """Runner that records concurrency and the memory held by running requests."""

    def __init__(self, hold=0.02):
        self.hold = hold
        self.lock = threading.Lock()
        self.running = 0
        self.peak = 0
        self.held = 0
        self.peak_held = 0
        self.order = []

    def __call__(self, req, device):
        need = serving.footprint_bytes(req)
        with self.lock:
            self.running += 1
            self.held += need
            self.peak = max(self.peak, self.running)
            self.peak_held = max(self.peak_held, self.held)
            self.order.append(req.rid)
        time.sleep(self.hold)
        with self.lock:
            self.running -= 1
            self.held -= need
        return torch.full((1, req.nt, 2), float(req.rid)), None, None


def test_footprint_grows_with_grid_and_gradient():
    small, big = _req(0, 128), _req(1, 512)
    assert serving.footprint_bytes(big) > 10 * serving.footprint_bytes(small)
    assert serving.footprint_bytes(_req(2, 256, gradient=True)) > serving.footprint_bytes(_req(3, 256))


def test_slots_and_budget_are_respected():
    fake = _Fake()
    reqs = [_req(k, 128 + 32 * (k % 5)) for k in range(24)]
    budget = 3 * max(serving.footprint_bytes(r) for r in reqs)
    with serving.SolveServer("cpu", slots=4, hbm_budget=budget, runner=fake) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=30) for f in futs]
    assert [r.rid for r in res] == list(range(24))
    assert all(float(r.receivers[0, 0, 0]) == r.rid for r in res)
    assert 1 < fake.peak <= 4
    assert fake.peak_held <= budget


def test_small_requests_pass_a_large_one_that_waits_for_memory():
    fake = _Fake(hold=0.05)
    big = _req(0, 512)
    smalls = [_req(k, 128) for k in range(1, 6)]
    budget = serving.footprint_bytes(big) + serving.footprint_bytes(smalls[0])
    with serving.SolveServer("cpu", slots=4, hbm_budget=budget, runner=fake) as srv:
        first = srv.submit(_req(99, 256))  # holds memory so the big one cannot start
        futs = [srv.submit(big)] + [srv.submit(r) for r in smalls]
        for f in [first] + futs:
            f.result(timeout=30)
    assert fake.order[0] == 99
    assert fake.order.index(0) > fake.order.index(1)  # a small one went ahead of the waiting big one
    assert fake.peak_held <= budget


def test_request_larger_than_budget_fails():
    with serving.SolveServer("cpu", slots=2, hbm_budget=1000, runner=_Fake()) as srv:
        with pytest.raises(MemoryError):
            srv.submit(_req(0)).result(timeout=10)


def test_serving_stream_mix():
    reqs = synthetic.serving_stream(200, rate=50.0, seed=5051, nt_scale=0.01)
    kinds = [("3d" if r.eps_r.ndim == 3 else "2d", r.gradient) for r in reqs]
    assert 0.7 < kinds.count(("2d", False)) / 200 < 0.9
    assert 0.08 < kinds.count(("2d", True)) / 200 < 0.22
    assert 0.01 < kinds.count(("3d", False)) / 200 < 0.1
    arrivals = [r.arrival for r in reqs]
    assert arrivals == sorted(arrivals)
    assert 2.5 < arrivals[-1] < 6.0  # 200 arrivals at 50/s


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix(cuda_device):
    reqs = synthetic.serving_stream(10, rate=200.0, seed=7, nt_scale=0.05)
    reqs = [r for r in reqs if r.eps_r.ndim == 2][:6]
    reqs[1].gradient = True
    alone = serving.run_sequential(reqs, cuda_device)["results"]
    mixed = serving.run_stream(reqs, cuda_device, slots=4, paced=False)["results"]
    for r in reqs:
        a, m = alone[r.rid], mixed[r.rid]
        assert torch.equal(a.receivers, m.receivers), r.rid
        if r.gradient:
            assert torch.equal(a.grad_eps_r, m.grad_eps_r)
            assert torch.equal(a.grad_sigma, m.grad_sigma)
