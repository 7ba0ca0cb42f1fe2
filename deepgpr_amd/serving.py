"""Request loop for forward and forward+gradient solves of mixed-size models
on one GPU.

Requests are admitted as slots and HBM free up (continuous batching): every
admitted request runs on its own HIP stream, so the kernels of requests in
flight together overlap on the device, and its state (fields, CPML
auxiliaries, histories) stays resident in HBM for the life of the request.
Admission prices each request with ``estimate_compute_memory`` (the byte
model of ``src/DeepGPR/solver/memory.py:30-268``); a request that does not fit
waits, and smaller requests behind it may go first within a bounded window.
A result does not depend on what else is in flight: the kernels are
deterministic and every request owns its arrays.

    server = SolveServer("cuda:0", slots=8)
    fut = server.submit(SolveRequest(...))
    result = fut.result()           # SolveResult: receivers (+ gradients), timings
    server.close()

``python -m deepgpr_amd.serving`` runs the mixed Poisson stream of
``synthetic.serving_stream`` through the server and one request at a time,
and prints one JSON line with p50/p99 latency and both throughputs.
"""
from __future__ import annotations

import collections
import concurrent.futures
import contextlib
import math
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import torch

from .solver.memory import estimate_compute_memory


@dataclass
class SolveRequest:
    """One solve: a physical model (``(nx, ny)`` or ``(nx, ny, nz)``), its
    sources and receivers, and ``compute`` options. ``gradient`` asks for the
    eps_r/sigma gradient of ``sum(receivers**2)``."""

    eps_r: torch.Tensor
    sigma: torch.Tensor
    source_amplitudes: torch.Tensor
    source_location: torch.Tensor
    receiver_location: torch.Tensor
    dx: float
    dt: float
    fdtd_order: int = 2
    pmlthick: int = 10
    gradient: bool = False
    options: Dict = field(default_factory=dict)
    rid: int = 0
    arrival: float = 0.0  # seconds after the stream starts (load generators)

    @property
    def nt(self) -> int:
        return int(self.source_amplitudes.shape[1])

    @property
    def extended_shape(self) -> List[int]:
        return [int(n) + 2 * self.pmlthick for n in self.eps_r.shape]

    def cell_updates(self) -> int:
        cells = math.prod(self.extended_shape) * int(self.source_location.shape[0])
        return cells * self.nt * (2 if self.gradient else 1)


@dataclass
class SolveResult:
    rid: int
    receivers: torch.Tensor
    grad_eps_r: Optional[torch.Tensor]
    grad_sigma: Optional[torch.Tensor]
    submitted: float
    started: float
    finished: float
    footprint: int

    @property
    def latency(self) -> float:
        return self.finished - self.submitted

    @property
    def queued(self) -> float:
        return self.started - self.submitted


def footprint_bytes(req: SolveRequest) -> int:
    """HBM bytes a request holds while it runs (fields, CPML, coefficients,
    history, receivers, adjoint state), plus 10 % for temporaries."""
    shape = req.extended_shape + ([1] if len(req.extended_shape) == 2 else [])
    opt = req.options
    storage = opt.get("wavefield_storage_dtype", torch.float32)
    name = {torch.float32: "float32", torch.float16: "float16", torch.bfloat16: "bfloat16"}[storage]
    est = estimate_compute_memory(*shape, req.nt, int(req.source_location.shape[0]), [req.pmlthick] * 6,
                                  mode=opt.get("mode", 2), sampling_interval=opt.get("model_gradient_sampling_interval", 1),
                                  storage=name, compression=opt.get("wavefield_compression", "none"),
                                  reconstruct_rhs=opt.get("wavefield_rhs_history", "auto") != "stored" and name == "float32",
                                  gradient=req.gradient, nrx=int(req.receiver_location.shape[1]))
    total = est["total"] - (0 if req.gradient else est["history"])  # forward-only solves keep no history
    return int(total * 1.1)


def solve(req: SolveRequest, device: torch.device):
    """Runs one request on the current stream; returns (receivers, grad_eps, grad_sigma)
    on the device. Forward-only requests keep no history."""
    from .solver.modeling import compute

    eps = req.eps_r.to(device, non_blocking=True)
    sig = req.sigma.to(device, non_blocking=True)
    amp = req.source_amplitudes.to(device, non_blocking=True)
    kw = dict(pmlthick=req.pmlthick, fdtd_order=req.fdtd_order, **req.options)
    if not req.gradient:
        with torch.no_grad():
            out = compute(device, req.dx, req.dt, amp, req.source_location, req.receiver_location, eps_r=eps,
                          sigma=sig, save_wavefield_history=False, **kw)
        return out[-1], None, None
    eps.requires_grad_(True)
    sig.requires_grad_(True)
    out = compute(device, req.dx, req.dt, amp, req.source_location, req.receiver_location, eps_r=eps, sigma=sig, **kw)
    rx = out[-1]
    g_eps, g_sig = torch.autograd.grad((rx * rx).sum(), [eps, sig])
    return rx.detach(), g_eps, g_sig


class SolveServer:
    """Continuous-batching request loop on one device (see the module docstring).

    ``slots``: requests in flight at once (one stream each). ``hbm_budget``:
    bytes the in-flight requests may hold together (default: ``hbm_fraction``
    of the free device memory when the server starts). ``window``: how far
    behind a request that does not fit the dispatcher looks for one that does.
    """

    def __init__(self, device="cuda:0", slots: int = 8, hbm_budget: Optional[int] = None,
                 hbm_fraction: float = 0.8, window: int = 8,
                 runner: Optional[Callable[[SolveRequest, torch.device], tuple]] = None):
        self.device = torch.device(device)
        if hbm_budget is None:
            hbm_budget = int(hbm_fraction * torch.cuda.mem_get_info(self.device)[0])
        self.budget = int(hbm_budget)
        self.slots = int(slots)
        self.window = int(window)
        self._runner = runner or solve
        self._pending: collections.deque = collections.deque()
        self._cv = threading.Condition()
        self._reserved = 0
        self._running = 0
        self._closing = False
        self._local = threading.local()
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=self.slots, thread_name_prefix="dgx-slot")
        self._dispatcher = threading.Thread(target=self._dispatch, name="dgx-dispatch", daemon=True)
        self._dispatcher.start()

    # ------------------------------------------------------------------ API
    def submit(self, req: SolveRequest) -> concurrent.futures.Future:
        fut: concurrent.futures.Future = concurrent.futures.Future()
        need = footprint_bytes(req)
        if need > self.budget:
            fut.set_exception(MemoryError(f"request {req.rid} needs {need} B, more than the HBM budget "
                                          f"{self.budget} B"))
            return fut
        with self._cv:
            if self._closing:
                raise RuntimeError("server is closed")
            self._pending.append((req, fut, need, time.perf_counter()))
            self._cv.notify_all()
        return fut

    def close(self, wait: bool = True) -> None:
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        self._dispatcher.join()
        self._pool.shutdown(wait=wait)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ internals
    def _pick(self):
        """Index of the next admissible pending request, or None."""
        if self._running >= self.slots:
            return None
        for k, (_, _, need, _) in enumerate(self._pending):
            if k >= self.window:
                break
            if self._reserved + need <= self.budget:
                return k
        return None

    def _dispatch(self) -> None:
        while True:
            with self._cv:
                while True:
                    k = self._pick()
                    if k is not None or (self._closing and not self._pending):
                        break
                    self._cv.wait()
                if k is None:
                    return
                item = self._pending[k]
                del self._pending[k]
                self._reserved += item[2]
                self._running += 1
            self._pool.submit(self._run, item)

    def _stream(self) -> torch.cuda.Stream:
        s = getattr(self._local, "stream", None)
        if s is None:
            s = torch.cuda.Stream(self.device)  # default priority
            self._local.stream = s
        return s

    def _run(self, item) -> None:
        req, fut, need, submitted = item
        started = time.perf_counter()
        try:
            on_gpu = self.device.type == "cuda"
            stream = self._stream() if on_gpu else None
            ctx = contextlib.ExitStack()
            if on_gpu:
                ctx.enter_context(torch.cuda.device(self.device))
                ctx.enter_context(torch.cuda.stream(stream))
            with ctx:
                rx, g_eps, g_sig = self._runner(req, self.device)
                rx = rx.to("cpu")
                g_eps = None if g_eps is None else g_eps.to("cpu")
                g_sig = None if g_sig is None else g_sig.to("cpu")
                if on_gpu:
                    stream.synchronize()
            res = SolveResult(req.rid, rx, g_eps, g_sig, submitted, started, time.perf_counter(), need)
            fut.set_result(res)
        except BaseException as exc:  # noqa: BLE001 - delivered to the caller
            fut.set_exception(exc)
        finally:
            with self._cv:
                self._reserved -= need
                self._running -= 1
                self._cv.notify_all()


def latency_stats(results: Sequence[SolveResult], wall: float) -> Dict:
    """p50/p99 latency (ms), requests/s and Gcell-updates/s of a finished run."""
    lat = sorted(r.latency for r in results)

    def pct(q):
        if not lat:
            return None
        idx = min(len(lat) - 1, max(0, math.ceil(q * len(lat)) - 1))
        return round(1e3 * lat[idx], 2)

    return {"requests": len(results), "p50_ms": pct(0.50), "p99_ms": pct(0.99),
            "requests_per_s": round(len(results) / wall, 3) if wall > 0 else None}


def run_stream(requests: Sequence[SolveRequest], device="cuda:0", slots: int = 8, paced: bool = True,
               hbm_budget: Optional[int] = None) -> Dict:
    """Feeds requests to a server at their arrival times (or all at once when
    ``paced`` is False); returns the stats and the results by request id."""
    results: Dict[int, SolveResult] = {}
    with SolveServer(device, slots=slots, hbm_budget=hbm_budget) as server:
        t0 = time.perf_counter()
        futs = []
        for req in requests:
            if paced:
                delay = req.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            futs.append(server.submit(req))
        for f in futs:
            r = f.result()
            results[r.rid] = r
        wall = time.perf_counter() - t0
    stats = latency_stats(list(results.values()), wall)
    stats["gcell_updates_per_s"] = round(sum(q.cell_updates() for q in requests) / wall / 1e9, 3)
    stats["wall_s"] = round(wall, 3)
    return {"stats": stats, "results": results}


def run_sequential(requests: Sequence[SolveRequest], device="cuda:0") -> Dict:
    """The same requests one at a time through plain compute(), back to back."""
    dev = torch.device(device)
    results: Dict[int, SolveResult] = {}
    t0 = time.perf_counter()
    for req in requests:
        s = time.perf_counter()
        rx, g_eps, g_sig = solve(req, dev)
        rx = rx.cpu()
        g_eps = None if g_eps is None else g_eps.cpu()
        g_sig = None if g_sig is None else g_sig.cpu()
        torch.cuda.synchronize(dev)
        e = time.perf_counter()
        results[req.rid] = SolveResult(req.rid, rx, g_eps, g_sig, s, s, e, footprint_bytes(req))
    wall = time.perf_counter() - t0
    stats = latency_stats(list(results.values()), wall)
    stats["gcell_updates_per_s"] = round(sum(q.cell_updates() for q in requests) / wall / 1e9, 3)
    stats["wall_s"] = round(wall, 3)
    return {"stats": stats, "results": results}


def _main() -> int:
    import argparse
    import json

    from . import synthetic

    ap = argparse.ArgumentParser(description="mixed Poisson stream: served vs one at a time")
    ap.add_argument("--requests", type=int, default=60)
    ap.add_argument("--rate", type=float, default=50.0, help="Poisson arrivals per second")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--seed", type=int, default=5051)
    ap.add_argument("--scale", type=float, default=1.0, help="multiplies every nt (shorter runs)")
    args = ap.parse_args()
    reqs = synthetic.serving_stream(args.requests, rate=args.rate, seed=args.seed, nt_scale=args.scale)
    dev = "cuda:0"
    # warm the library and the allocator on one small request
    warm = synthetic.serving_stream(1, rate=args.rate, seed=args.seed + 1, nt_scale=0.05)
    run_sequential(warm, dev)
    seq = run_sequential(reqs, dev)
    paced = run_stream(reqs, dev, slots=args.slots, paced=True)
    burst = run_stream(reqs, dev, slots=args.slots, paced=False)
    worst = 0.0
    for rid, r in burst["results"].items():
        ref = seq["results"][rid]
        worst = max(worst, float((r.receivers - ref.receivers).abs().max()))
    line = {"metric": "serve p50/p99 ms vs plain solve", "requests": len(reqs), "rate_per_s": args.rate,
            "slots": args.slots, "nt_scale": args.scale, "sequential": seq["stats"], "served_paced": paced["stats"],
            "served_saturated": burst["stats"],
            "throughput_ratio": round(burst["stats"]["gcell_updates_per_s"] / seq["stats"]["gcell_updates_per_s"], 3),
            "max_abs_receiver_diff_vs_alone": worst}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
