"""Request loop for forward and forward+gradient solves of mixed-size models
on one GPU.

Requests are admitted as slots and HBM free up (continuous batching). 2D
Ez-TM requests join one running batch (``batch.BatchSolver``): every time
step of every such request in flight is one kernel launch per FDTD order and
phase, and a request joins or leaves between two groups of steps. By default
they do so only while no request of another kind is waiting or running, and
take a stream otherwise (``batch_policy``). Other requests (3D, INT8
histories, ...) each run on a HIP stream of their own, so their kernels
overlap on the device. A request's state (fields, CPML
auxiliaries, histories) stays resident in HBM for the life of the request.
Admission prices each request with ``solve_footprint`` (a byte model in the
manner of ``src/DeepGPR/solver/memory.py:30-246``); a request that does not fit
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

from .utils.devices import resolve_device
from .solver.memory import solve_footprint


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
    est = solve_footprint(*shape, req.nt, int(req.source_location.shape[0]), [req.pmlthick] * 6,
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

    ``slots``: requests in flight at once outside the batch (one stream each).
    ``batch_slots``: 2D Ez-TM requests in the running batch at once (0 turns
    the batch off; it is off with a custom ``runner`` or on the CPU).
    ``batch_chunk``: time steps between two points where requests join or
    leave the batch. ``batch_policy``: ``"alone"`` (default) sends a
    batchable request to the batch only while no request of another kind is
    waiting or running, and to a stream otherwise; ``"always"`` sends every
    batchable request to the batch. The batch is one chain of small dependent
    kernels: beside the large kernels of other requests each of its steps
    waits for free CUs, while on separate streams the small requests spread
    that wait over several chains. ``batch_solver``: factory ``f(device)`` of
    the object that runs the batch (default ``batch.BatchSolver``). ``hbm_budget``: bytes the in-flight requests may hold
    together (default: ``hbm_fraction`` of the free device memory when the
    server starts). ``window``: how far behind a request that does not fit the
    dispatcher looks for one that does.
    """

    def __init__(self, device="cuda:0", slots: int = 8, hbm_budget: Optional[int] = None,
                 hbm_fraction: float = 0.8, window: int = 8,
                 runner: Optional[Callable[[SolveRequest, torch.device], tuple]] = None,
                 batch_slots: int = 64, batch_chunk: int = 32, batch_solver: Optional[Callable] = None,
                 batch_policy: str = "alone"):
        if batch_policy not in ("alone", "always"):
            raise ValueError(f"batch_policy must be 'alone' or 'always', not {batch_policy!r}")
        self.device = resolve_device(device)
        if hbm_budget is None:
            hbm_budget = int(hbm_fraction * torch.cuda.mem_get_info(self.device)[0])
        self.budget = int(hbm_budget)
        self.slots = int(slots)
        self.window = int(window)
        self._runner = runner or solve
        self._batch_solver = batch_solver
        use_batch = batch_slots > 0 and (batch_solver is not None or (runner is None and self.device.type == "cuda"))
        self.batch_slots = int(batch_slots) if use_batch else 0
        self.batch_chunk = int(batch_chunk)
        self.batch_policy = batch_policy
        self._others = 0  # requests the batch cannot serve, waiting or running
        self._pending: collections.deque = collections.deque()
        self._cv = threading.Condition()
        self._reserved = 0
        self._running = 0
        self._batched = 0
        self._inbox: List = []
        self._closing = False
        self._dispatch_over = False
        self._local = threading.local()
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=self.slots, thread_name_prefix="dgx-slot")
        self._batcher = None
        if self.batch_slots:
            self._batcher = threading.Thread(target=self._batch_loop, name="dgx-batch", daemon=True)
            self._batcher.start()
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
            self._others += 0 if self._batchable(req) else 1
            self._pending.append((req, fut, need, time.perf_counter()))
            self._cv.notify_all()
        return fut

    def close(self, wait: bool = True) -> None:
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        self._dispatcher.join()
        if self._batcher is not None:
            self._batcher.join()
        self._pool.shutdown(wait=wait)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ internals
    @staticmethod
    def _batchable(req: SolveRequest) -> bool:
        from .batch import batchable
        return batchable(req)

    def _to_batch(self, req: SolveRequest) -> bool:
        if self.batch_slots <= 0 or not self._batchable(req):
            return False
        return self.batch_policy == "always" or self._others == 0

    def _pick(self):
        """(index, batched) of the next admissible pending request, or None."""
        for k, (req, _, need, _) in enumerate(self._pending):
            if k >= self.window:
                break
            if self._reserved + need > self.budget:
                continue
            batched = self._to_batch(req)
            if (self._batched < self.batch_slots) if batched else (self._running < self.slots):
                return k, batched
        return None

    def _dispatch(self) -> None:
        while True:
            with self._cv:
                while True:
                    pick = self._pick()
                    if pick is not None or (self._closing and not self._pending):
                        break
                    self._cv.wait()
                if pick is None:
                    self._dispatch_over = True
                    self._cv.notify_all()
                    return
                k, batched = pick
                item = self._pending[k]
                del self._pending[k]
                self._reserved += item[2]
                if batched:
                    self._batched += 1
                    self._inbox.append(item)
                    self._cv.notify_all()
                    continue
                self._running += 1
            self._pool.submit(self._run, item)

    def _release(self, need: int, batched: bool, other: bool = False) -> None:
        with self._cv:
            self._reserved -= need
            if batched:
                self._batched -= 1
            else:
                self._running -= 1
            self._others -= 1 if other else 0
            self._cv.notify_all()

    def _batch_loop(self) -> None:
        """Owns the running batch: adds what the dispatcher hands over, steps
        every instance ``batch_chunk`` time steps at a time and delivers the
        results of finished instances."""
        from .batch import BatchSolver

        live: Dict[int, tuple] = {}
        on_gpu = self.device.type == "cuda"
        with torch.cuda.device(self.device) if on_gpu else contextlib.nullcontext():
            solver = None
            try:
                solver = (self._batch_solver or BatchSolver)(self.device)
                while True:
                    with self._cv:
                        while not self._inbox and not live and not self._dispatch_over:
                            self._cv.wait()
                        if not self._inbox and not live:
                            return
                        new, self._inbox = self._inbox, []
                    for item in new:  # prepared on the solver's thread while the batch steps
                        live[solver.submit(item[0])] = (item, time.perf_counter())
                    ready = solver.step(self.batch_chunk if solver.active else 0)
                    if not ready and live and not solver.active:
                        if solver.pending:  # only preparations are left: wait for them or new work
                            with self._cv:
                                if not self._inbox:
                                    self._cv.wait(timeout=0.0005)
                        else:
                            solver.synchronize()  # only completion events are left
                    for nid in ready:
                        (req, fut, need, submitted), started = live.pop(nid)
                        try:
                            rx, g_eps, g_sig = solver.result(nid)
                            fut.set_result(SolveResult(req.rid, rx, g_eps, g_sig, submitted, started,
                                                       time.perf_counter(), need))
                        except BaseException as exc:  # noqa: BLE001 - delivered to the caller
                            fut.set_exception(exc)
                        self._release(need, True)
            except BaseException as exc:  # noqa: BLE001 - every waiting caller hears of it
                for (req, fut, need, _), _ in live.values():
                    if not fut.done():
                        fut.set_exception(exc)
                    self._release(need, True)
                with self._cv:
                    stranded, self._inbox = self._inbox, []
                    self.batch_slots = 0  # later requests take the stream path
                    self._cv.notify_all()
                for req, fut, need, _ in stranded:
                    fut.set_exception(exc)
                    self._release(need, True)
            finally:
                if solver is not None:
                    solver.close()

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
            self._release(need, False, not self._batchable(req))


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
               hbm_budget: Optional[int] = None, batch_slots: int = 64, batch_chunk: int = 32,
               batch_policy: str = "alone") -> Dict:
    """Feeds requests to a server at their arrival times (or all at once when
    ``paced`` is False); returns the stats and the results by request id."""
    results: Dict[int, SolveResult] = {}
    with SolveServer(device, slots=slots, hbm_budget=hbm_budget, batch_slots=batch_slots,
                     batch_chunk=batch_chunk, batch_policy=batch_policy) as server:
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
    dev = resolve_device(device)
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
    ap.add_argument("--batch-slots", type=int, default=64, help="2D requests in the running batch (0: off)")
    ap.add_argument("--batch-chunk", type=int, default=32, help="time steps between joins")
    ap.add_argument("--batch-policy", default="alone", choices=("alone", "always"),
                    help="alone: 2D requests join the batch only while no other kind is in the server")
    ap.add_argument("--streams-only", action="store_true", help="also run the server with the batch off")
    args = ap.parse_args()
    reqs = synthetic.serving_stream(args.requests, rate=args.rate, seed=args.seed, nt_scale=args.scale)
    dev = "cuda:0"
    # warm the library and the allocator on one small request
    warm = synthetic.serving_stream(1, rate=args.rate, seed=args.seed + 1, nt_scale=0.05)
    run_sequential(warm, dev)
    seq = run_sequential(reqs, dev)
    kw = dict(slots=args.slots, batch_slots=args.batch_slots, batch_chunk=args.batch_chunk,
              batch_policy=args.batch_policy)
    paced = run_stream(reqs, dev, paced=True, **kw)
    burst = run_stream(reqs, dev, paced=False, **kw)
    worst = 0.0
    for rid, r in burst["results"].items():
        ref = seq["results"][rid]
        worst = max(worst, float((r.receivers - ref.receivers).abs().max()))
    line = {"metric": "serve p50/p99 ms vs plain solve", "requests": len(reqs), "rate_per_s": args.rate,
            "slots": args.slots, "batch_slots": args.batch_slots, "batch_chunk": args.batch_chunk,
            "batch_policy": args.batch_policy,
            "nt_scale": args.scale, "sequential": seq["stats"], "served_paced": paced["stats"],
            "served_saturated": burst["stats"],
            "throughput_ratio": round(burst["stats"]["gcell_updates_per_s"] / seq["stats"]["gcell_updates_per_s"], 3),
            "max_abs_receiver_diff_vs_alone": worst}
    if args.streams_only:
        alt = run_stream(reqs, dev, paced=False, slots=args.slots, batch_slots=0)
        line["served_saturated_streams_only"] = alt["stats"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
