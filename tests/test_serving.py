"""Admission, slot and ordering behaviour of the request loop (CPU, fake solver),
and GPU checks that a result is the same alone and in the mix."""
import concurrent.futures
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


class _FakeBatch:
    """Batch runner: an instance finishes once ``step`` has advanced it nt steps."""

    made = []

    def __init__(self, device, fail_rid=None):
        self.left, self.reqs, self.next, self.peak, self.closed = {}, {}, 0, 0, False
        self.fail_rid = fail_rid
        self.failed = set()
        self.seen = []
        _FakeBatch.made.append(self)

    pending = 0

    def submit(self, req):
        nid, self.next = self.next, self.next + 1
        if req.rid == self.fail_rid:
            self.failed.add(nid)
            return nid
        self.left[nid], self.reqs[nid] = req.nt, req
        self.peak = max(self.peak, len(self.left))
        self.seen.append(req.rid)
        return nid

    def step(self, steps):
        time.sleep(0.002 if steps else 0.0)  # a group of steps takes device time
        ready = sorted(self.failed)
        self.failed.clear()
        ready += [n for n in self.left if self.left[n] <= steps]
        for n in list(self.left):
            self.left[n] -= steps
            if self.left[n] <= 0:
                del self.left[n]
        return ready

    @property
    def active(self):
        return len(self.left)

    def synchronize(self):
        pass

    def result(self, nid):
        if nid not in self.reqs:
            raise ValueError("cannot batch this one")
        req = self.reqs.pop(nid)
        return torch.full((1, req.nt, 2), float(req.rid)), None, None

    def close(self):
        self.closed = True


def _req3d(rid, side=32, nt=40):
    eps = torch.full((side, side, side), 4.0)
    wave = torch.zeros((1, nt, 1))
    src = torch.tensor([[[5, side // 2, side // 2]]])
    rx = torch.tensor([[[5, 3, 3]]])
    return serving.SolveRequest(eps, torch.zeros_like(eps), wave, src, rx, 0.02, 2.5e-11, rid=rid)


@pytest.mark.parametrize("fail_rid", [None, 4])
def test_batchable_requests_share_the_batch_and_the_rest_take_streams(fail_rid):
    _FakeBatch.made.clear()
    fake = _Fake(hold=0.01)
    reqs = []
    for k in range(20):
        r = _req3d(k) if k % 5 == 0 else _req(k, 128, nt=30 + 17 * k)
        if k == 7:
            r.options = {"model_gradient_sampling_interval": 3}  # not batchable: stream path
        reqs.append(r)
    with serving.SolveServer("cpu", slots=2, hbm_budget=1 << 40, runner=fake, batch_slots=6, batch_chunk=8,
                             batch_solver=lambda dev: _FakeBatch(dev, fail_rid), batch_policy="always") as srv:
        futs = [srv.submit(r) for r in reqs]
        done = concurrent.futures.wait(futs, timeout=30)
        assert not done.not_done
    (batch,) = _FakeBatch.made
    assert batch.closed and 1 < batch.peak <= 6
    streamed = sorted(fake.order)
    assert streamed == [0, 5, 7, 10, 15]
    assert sorted(batch.seen + streamed + ([fail_rid] if fail_rid else [])) == list(range(20))
    for r, f in zip(reqs, futs):
        if r.rid == fail_rid:
            with pytest.raises(ValueError):
                f.result()
            continue
        res = f.result()
        assert res.rid == r.rid and float(res.receivers[0, 0, 0]) == r.rid
        assert res.receivers.shape[1] == r.nt


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


def test_batch_policy_alone_streams_2d_requests_beside_others():
    """With the default policy a 2D request takes a stream while a 3D request
    is waiting or running, and joins the batch once the server holds none."""
    _FakeBatch.made.clear()
    fake = _Fake(hold=0.05)
    first = [_req3d(0)] + [_req(k, 128, nt=40) for k in range(1, 5)]
    with serving.SolveServer("cpu", slots=8, hbm_budget=1 << 40, runner=fake, batch_slots=6, batch_chunk=8,
                             batch_solver=lambda dev: _FakeBatch(dev)) as srv:
        assert srv.batch_policy == "alone"
        futs = [srv.submit(r) for r in first]
        assert not concurrent.futures.wait(futs, timeout=30).not_done
        later = [_req(k, 128, nt=40) for k in range(5, 9)]
        futs += [srv.submit(r) for r in later]
        assert not concurrent.futures.wait(futs, timeout=30).not_done
    (batch,) = _FakeBatch.made
    assert sorted(fake.order) == [0, 1, 2, 3, 4]
    assert sorted(batch.seen) == [5, 6, 7, 8]
    for r, f in zip(first + later, futs):
        assert float(f.result().receivers[0, 0, 0]) == r.rid
    with pytest.raises(ValueError):
        serving.SolveServer("cpu", runner=fake, batch_policy="sometimes")


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
@pytest.mark.parametrize("batch_slots,policy", [(0, "alone"), (3, "always"), (3, "alone")])
def test_results_same_alone_and_in_the_mix(cuda_device, batch_slots, policy):
    reqs = synthetic.serving_stream(10, rate=200.0, seed=7, nt_scale=0.05)
    reqs = [r for r in reqs if r.eps_r.ndim == 2][:6]
    reqs[1].gradient = True
    reqs[2].options = {"mode": 2, "model_gradient_sampling_interval": 2}  # not batchable
    reqs[2].gradient = True
    reqs.append([r for r in synthetic.serving_stream(200, seed=8, nt_scale=0.02) if r.eps_r.ndim == 3][0])
    reqs[-1].rid = 99
    alone = serving.run_sequential(reqs, cuda_device)["results"]
    mixed = serving.run_stream(reqs, cuda_device, slots=2, paced=False, batch_slots=batch_slots,
                               batch_chunk=5, batch_policy=policy)["results"]
    for r in reqs:
        a, m = alone[r.rid], mixed[r.rid]
        assert torch.equal(a.receivers, m.receivers), r.rid
        if r.gradient:
            assert torch.equal(a.grad_eps_r, m.grad_eps_r)
            assert torch.equal(a.grad_sigma, m.grad_sigma)
