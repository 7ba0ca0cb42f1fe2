"""Batched 2D Ez-TM solves: many independent forward or forward+gradient
solves of different grids, FDTD orders and numbers of time steps advance
together, one kernel launch per (order, phase) and time step (``dgx_batch_*``
in the C ABI; the reference batches only the shots of one model,
``deepgpr.cu:618, 3493``).

    solver = BatchSolver("cuda:0")
    ids = [solver.submit(req) for req in requests]  # serving.SolveRequest, 2D
    solver.run()                                    # until every instance finished
    rx, g_eps, g_sig = solver.result(ids[0])        # physical-grid gradients

Each instance is prepared exactly as ``compute`` prepares it
(``prepare_compute``) and advanced by the same fused 2D TM kernel bodies, so
its receivers and gradients equal those of ``compute`` for the same request.
A gradient instance asks for the eps_r / sigma gradient of
``sum(receivers**2)`` and keeps an E-only history at sampling interval 1.

``python -m deepgpr_amd.batch`` runs the 64-instance batch of
``synthetic.batch_instances`` one at a time through ``compute`` and as one
batch (prepared on a background thread, longest first, joining while the
batch steps), checks that the results agree and prints one JSON line.
"""
from __future__ import annotations

import collections
import ctypes
import itertools
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import torch

from .native.bridge import build_solve
from .native.loader import require_library
from .solver.autograd import forward_native_args
from .solver.modeling import prepare_compute
from .utils.devices import resolve_device
from .utils.exceptions import NativeLibraryError


@dataclass
class _Instance:
    rid: int
    gradient: bool
    receivers: torch.Tensor
    eps: Optional[torch.Tensor]
    sig: Optional[torch.Tensor]
    eps_ext: Optional[torch.Tensor]
    sig_ext: Optional[torch.Tensor]
    g_eps: Optional[torch.Tensor]
    g_sig: Optional[torch.Tensor]
    keep: List = field(default_factory=list)  # everything the native instance points into


def batchable(req) -> bool:
    """Requests the batch runtime serves with the result ``compute`` gives:
    2D models in the Ez-TM mode, no INT8 history and, for gradients, the
    E-only history at sampling interval 1 that ``compute`` picks for them by
    default."""
    o = req.options
    rhs = o.get("wavefield_rhs_history", "auto")
    lossless = o.get("wavefield_storage_dtype", torch.float32) == torch.float32
    return (req.eps_r.ndim == 2 and o.get("mode", 2) == 2 and o.get("wavefield_compression", "none") == "none"
            and o.get("model_gradient_sampling_interval", 1) == 1
            and (rhs == "reconstructed" or (rhs == "auto" and lossless)))


@dataclass
class _Prepared:
    req: object
    fwd: object
    grad: object
    inst: _Instance


class BatchSolver:
    """Drives one ``dgx_batch`` on its own stream from one host thread.

    ``add`` prepares a request on the calling thread and puts it in the
    batch. ``submit`` hands it to a background thread that prepares it while
    the batch steps; it joins at a later ``step``. Both return an id that
    ``step`` reports once the result is complete and ``result`` reads."""

    def __init__(self, device="cuda:0", stream: Optional[torch.cuda.Stream] = None, lookahead: int = 2):
        self.device = resolve_device(device)
        self.lookahead = int(lookahead)  # step() calls the host may run ahead of the device
        self._marks: collections.deque = collections.deque()
        self._lib = require_library(self.device)
        self.stream = stream or torch.cuda.Stream(self.device)  # default priority; time steps
        # requests are prepared and results read back on a second stream, so
        # that their host synchronisations do not wait for queued time steps
        self.side = torch.cuda.Stream(self.device)
        self._handle = self._lib.dgx_batch_create(ctypes.c_void_p(self.stream.cuda_stream))
        if not self._handle:
            raise NativeLibraryError(self._error())
        self._ids = itertools.count(1)
        self._running: Dict[int, tuple] = {}  # id -> (native id, _Instance)
        self._done: Dict[int, object] = {}    # id -> _Instance or the exception that ended it
        self._pending = 0                     # submitted, not yet in the batch
        self._inbox: queue.Queue = queue.Queue()
        self._prepared: collections.deque = collections.deque()
        self._worker: Optional[threading.Thread] = None

    def _error(self) -> str:
        msg = self._lib.dgx_last_error()
        return msg.decode() if isinstance(msg, bytes) else str(msg)

    # ---------------------------------------------------------- preparation
    def _prepare(self, req) -> _Prepared:
        if not batchable(req):
            raise ValueError("batched solves are 2D Ez-TM solves (mode 2, no INT8 history)")
        dev = self.device
        with torch.cuda.device(dev), torch.cuda.stream(self.side):
            eps = req.eps_r.to(dev)
            sig = req.sigma.to(dev)
            if req.gradient:
                eps.requires_grad_(True)
                sig.requires_grad_(True)
            amp = req.source_amplitudes.to(dev)
            kw = dict(req.options)
            kw.setdefault("mode", 2)
            if req.gradient:
                kw["wavefield_rhs_history"] = "reconstructed"  # E-only history
                kw["model_gradient_sampling_interval"] = 1
            p = prepare_compute(dev, req.dx, req.dt, amp, req.source_location, req.receiver_location, eps_r=eps,
                                sigma=sig, pmlthick=req.pmlthick, fdtd_order=req.fdtd_order,
                                save_wavefield_history=req.gradient, **kw)
            if not p.config.tm2d_fast_path:
                raise ValueError("batched solves need the 2D Ez-TM path (Ez source and receivers)")
            states = (*p.electric, *p.magnetic, *p.phi)
            args, out = forward_native_args(p.config, p.tensors, p.waveforms.contiguous(), states, req.gradient)
            fwd = build_solve(args)
            grad = None
            g_eps = g_sig = None
            if req.gradient:
                cells = (p.config.nx, p.config.ny, p.config.nz)
                g_eps = torch.zeros(cells, device=dev, dtype=p.config.dtype)
                g_sig = torch.zeros(cells, device=dev, dtype=p.config.dtype)
                grad = build_solve({"grad_eps_r": g_eps, "grad_sigma": g_sig, "eps_r_requires_grad": 1,
                                    "sigma_requires_grad": 1})
        inst = _Instance(req.rid, req.gradient, out["receivers"], eps, sig, p.eps_ext, p.sig_ext, g_eps, g_sig,
                         keep=[args, out, p, fwd, grad, amp])
        return _Prepared(req, fwd, grad, inst)

    def _commit(self, pid: int, prep: _Prepared) -> None:
        with torch.cuda.device(self.device):
            nid = self._lib.dgx_batch_add(self._handle, ctypes.byref(prep.fwd),
                                          ctypes.byref(prep.grad) if prep.grad is not None else None,
                                          ctypes.c_void_p(self.side.cuda_stream))
        if nid < 0:
            raise NativeLibraryError(self._error())
        self._running[pid] = (nid, prep.inst)

    def _work(self) -> None:
        while True:
            item = self._inbox.get()
            if item is None:
                return
            pid, req = item
            try:
                self._prepared.append((pid, self._prepare(req)))
            except BaseException as exc:  # noqa: BLE001 - reported through result()
                self._prepared.append((pid, exc))

    # ------------------------------------------------------------------ API
    def add(self, req) -> int:
        """Prepares a request on this thread and adds it to the batch; returns its id."""
        pid = next(self._ids)
        self._commit(pid, self._prepare(req))
        return pid

    def submit(self, req) -> int:
        """Queues a request for preparation on the background thread; it joins
        the batch at a later step(). Returns its id."""
        if self._worker is None:
            self._worker = threading.Thread(target=self._work, name="dgx-batch-prepare", daemon=True)
            self._worker.start()
        pid = next(self._ids)
        self._pending += 1
        self._inbox.put((pid, req))
        return pid

    @property
    def pending(self) -> int:
        """Submitted requests not yet in the batch."""
        return self._pending

    def step(self, steps: int = 1) -> List[int]:
        """Adds the requests prepared since the last call, advances every
        instance by up to ``steps`` time steps and returns the ids whose
        results are complete (or whose preparation failed)."""
        ready = []
        while self._prepared:
            pid, prep = self._prepared.popleft()
            self._pending -= 1
            if isinstance(prep, BaseException):
                self._done[pid] = prep
                ready.append(pid)
                continue
            try:
                self._commit(pid, prep)
            except NativeLibraryError as exc:
                self._done[pid] = exc
                ready.append(pid)
        n = self._lib.dgx_batch_step(self._handle, int(steps))
        if n < 0:
            raise NativeLibraryError(self._error())
        if steps > 0 and self.active:
            # bound the queued work so that an instance added next joins soon
            mark = torch.cuda.Event()
            mark.record(self.stream)
            self._marks.append(mark)
            while len(self._marks) > self.lookahead:
                self._marks.popleft().synchronize()
        for pid in list(self._running):
            nid, inst = self._running[pid]
            state = self._lib.dgx_batch_poll(self._handle, nid)
            if state < 0:
                raise NativeLibraryError(self._error())
            if state == 1:
                del self._running[pid]
                self._done[pid] = inst
                ready.append(pid)
        return ready

    def synchronize(self) -> None:
        self.stream.synchronize()

    @property
    def active(self) -> int:
        return int(self._lib.dgx_batch_active(self._handle))

    def run(self, chunk: int = 64) -> List[int]:
        """Steps until every instance has finished and every submitted request
        has been added; returns the finished ids."""
        ready = []
        while self._running or self._pending:
            ready += self.step(chunk if self.active else 0)
            if self._running and not self.active:
                self.synchronize()  # only completion events are left
            elif self._pending and not self._prepared and not self.active:
                time.sleep(0.0005)  # the preparation thread is working
        return ready

    def result(self, pid: int):
        """(receivers, grad_eps_r, grad_sigma) of a finished instance, on the
        CPU; gradients are folded onto the physical model (None for a
        forward-only instance). Raises what ended a failed one."""
        inst = self._done.pop(pid)  # its completion event has passed
        if isinstance(inst, BaseException):
            raise inst
        with torch.cuda.device(self.device), torch.cuda.stream(self.side):
            g_eps = g_sig = None
            if inst.gradient:
                g_eps, g_sig = torch.autograd.grad([inst.eps_ext, inst.sig_ext], [inst.eps, inst.sig],
                                                   [inst.g_eps, inst.g_sig])
                g_eps, g_sig = g_eps.to("cpu"), g_sig.to("cpu")
            rx = inst.receivers.to("cpu")
        return rx, g_eps, g_sig

    def close(self) -> None:
        if self._worker is not None:
            self._inbox.put(None)
            self._worker.join()
            self._worker = None
        if self._handle:
            self._lib.dgx_batch_destroy(self._handle)
            self._handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _main() -> int:
    import argparse
    import json

    from . import synthetic
    from .serving import run_sequential

    ap = argparse.ArgumentParser(description="64 uneven 2D forward+adjoint instances: one at a time vs batched")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--seed", type=int, default=4041)
    ap.add_argument("--scale", type=float, default=1.0, help="multiplies every nt (shorter runs)")
    ap.add_argument("--chunk", type=int, default=64, help="time steps per dgx_batch_step call")
    args = ap.parse_args()
    reqs = synthetic.batch_instances(args.instances, seed=args.seed, nt_scale=args.scale)
    dev = torch.device("cuda:0")
    # warm the library, the allocator and the tile tuner (one probe per shape)
    run_sequential(reqs, dev)
    seq = run_sequential(reqs, dev)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    with BatchSolver(dev) as solver:
        # longest first; a background thread prepares requests while the batch steps
        ids = {solver.submit(r): r.rid for r in sorted(reqs, key=lambda q: (2 if q.gradient else 1) * q.nt,
                                                         reverse=True)}
        solver.run(args.chunk)
        results = {ids[n]: solver.result(n) for n in ids}
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    work = sum(r.cell_updates() for r in reqs)
    rx_diff = max(float((results[r.rid][0] - seq["results"][r.rid].receivers).abs().max()) for r in reqs)
    g_rel = 0.0
    for r in reqs:
        if r.gradient:
            a, b = results[r.rid][1], seq["results"][r.rid].grad_eps_r
            g_rel = max(g_rel, float((a - b).norm() / b.norm().clamp_min(1e-30)))
    line = {"metric": "batched uneven instances vs one at a time", "instances": len(reqs), "nt_scale": args.scale,
            "sequential_s": seq["stats"]["wall_s"], "batched_s": round(wall, 3),
            "sequential_gcell_per_s": seq["stats"]["gcell_updates_per_s"],
            "batched_gcell_per_s": round(work / wall / 1e9, 3),
            "speedup": round(seq["stats"]["wall_s"] / wall, 3), "max_abs_receiver_diff": rx_diff,
            "max_grad_eps_rel_diff": g_rel}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
