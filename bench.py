#!/usr/bin/env python3
"""Benchmark of the FDTD hot path: forward + exact adjoint (material gradient)
of a seeded synthetic workload on one MI355X, through the public ``compute``
API and ``loss.backward()``.

One step = one compute() call and its backward: forward + adjoint of one
batch of shots, the batches taken in turn. A batch is every shot when the
whole solve fits in HBM (C1, C2) and one shot on the 3D model (one shot's
FP32 E+R history at nt 2000 is 223 GB). Prints one JSON line with the
whole-job throughput in Gcell-updates/s (extended grid cells incl. CPML x
time steps x shots x 2 directions), the roofline of the dominant kernel (HIP
events around each of its launches, on the launch stream) and the reference
CPU backend timed on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=("c1", "c2", "c3"))
    ap.add_argument("--nshot", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--storage", default="float32", choices=("float32", "bfloat16", "float16", "int8"))
    ap.add_argument("--rhs", default="auto", choices=("auto", "stored", "reconstructed"))
    ap.add_argument("--accumulator", default="float64", choices=("float32", "float64"))
    ap.add_argument("--shots-per-call", type=int, default=0,
                    help="shots per compute() call (0: all that fit in free HBM)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--cpu-baseline-child", default=None, help=argparse.SUPPRESS)
    return ap.parse_args()


def _workload(args):
    from deepgpr_amd import synthetic

    kw = {}
    if args.nshot is not None:
        kw["nshot"] = args.nshot
    if args.nt is not None:
        kw["nt"] = args.nt
    return synthetic.WORKLOADS[args.workload](**kw) if args.workload != "c1" else synthetic.c1_small_2d()


def _compute_kwargs(w, args):
    import torch

    kw = dict(pmlthick=w.pml, fdtd_order=w.fdtd_order, mode=w.mode, model_gradient_sampling_interval=w.sampling_interval,
              wavefield_rhs_history=args.rhs, gradient_accumulator=args.accumulator, **w.kwargs)
    if args.storage == "int8":
        kw["wavefield_compression"] = "int8"
    else:
        kw["wavefield_storage_dtype"] = {"float32": torch.float32, "bfloat16": torch.bfloat16,
                                         "float16": torch.float16}[args.storage]
    return kw


# ----------------------------------------------------------------------------- byte models


def _psi_elems(w):
    """Active CPML auxiliary elements per shot (E and H), 2D TM or 3D."""
    nx, ny = w.physical_shape[0] + 2 * w.pml, w.physical_shape[1] + 2 * w.pml
    p = w.pml
    if len(w.physical_shape) == 2:
        # x faces: E phi2 (p+1, ny+1), H phi1 (p, ny+1); y faces: E phi2 (nx+1, p+1), H phi1 (nx+1, p)
        e = 2 * ((p + 1) * (ny + 1) + (nx + 1) * (p + 1))
        h = 2 * (p * (ny + 1) + (nx + 1) * p)
        return e, h
    nz = w.physical_shape[2] + 2 * p
    e = h = 0
    for a, b in ((ny, nz), (nx, nz), (nx, ny)):
        e += 2 * 2 * (p + 1) * (a + 1) * (b + 1)
        h += 2 * 2 * p * (a + 1) * (b + 1)
    return e, h


def _history_terms(w, args):
    """(components, bytes per stored element, frames per saved step) of the history."""
    comps = 3 if w.mode == 3 else 1
    two_d = len(w.physical_shape) == 2
    tile = 64  # INT8 scale tile: 8x8x1 in 2D, 4x4x4 in 3D
    hist_b = {"float32": 4.0, "bfloat16": 2.0, "float16": 2.0, "int8": 1.0 + 4.0 / tile}[args.storage]
    e_only = args.rhs == "reconstructed" or (args.rhs == "auto" and args.storage == "float32"
                                             and w.sampling_interval == 1)
    return comps, hist_b, (1 if e_only else 2)


def model_bytes_per_cell(w, args):
    """Algorithmic HBM bytes per cell-update (forward, adjoint) of a one-sweep
    leapfrog step: fields read and written plus ca and cb (2D TM 32 B, 3D 56 B),
    8 B per active CPML auxiliary element, the history written (forward) and
    read (adjoint, E^n and E^(n+1) or R^n), and the two gradient accumulators
    read and written once per saved step."""
    two_d = len(w.physical_shape) == 2
    core = 32.0 if two_d else 56.0
    pe, ph = _psi_elems(w)
    cpml = 8.0 * (pe + ph) / w.extended_cells()
    comps, hist_b, frames = _history_terms(w, args)
    S = w.sampling_interval
    acc = 8.0 if args.accumulator == "float64" else 4.0
    fwd = core + cpml + comps * hist_b * frames / S
    adj = core + cpml + 2 * comps * hist_b / S + 2 * 2 * acc / S
    return fwd, adj


def kernel_bytes(kernel: str, w, args) -> float:
    """Algorithmic bytes of ONE launch of a kernel (one time step, all shots of
    one call). The fused 2D TM kernels carry the per-cell model above times
    the cells they advance; the gradient accumulators belong to the reduction
    kernel while it is separate. The split kernels carry their own share of
    the compulsory loads/stores."""
    shots = args.shots_per_call or w.source_location.shape[0]
    ext = w.extended_cells()
    cells = ext * shots
    two_d = len(w.physical_shape) == 2
    pe, ph = (x * shots for x in _psi_elems(w))
    comps, hist_b, frames = _history_terms(w, args)
    fwd, adj = model_bytes_per_cell(w, args)
    acc_b = 2 * 2 * (8.0 if args.accumulator == "float64" else 4.0) / w.sampling_interval
    if kernel == "tm_step_fwd":
        return fwd * cells
    if kernel == "tm_step_adj":
        return (adj - acc_b) * cells
    if kernel == "tm_grad_reduce":
        return acc_b * w.sampling_interval * ext  # one launch per saved step, accumulators of all shots
    if kernel == "fwd_h":
        core = 24 if two_d else 40      # E read, H read+write, ch_curl
        return cells * core + 8 * ph
    if kernel == "fwd_e":
        core = 24 if two_d else 48      # H read, E read+write, ca, ce_curl (+ce_rhs in 3D)
        return cells * (core + comps * hist_b * frames / w.sampling_interval) + 8 * pe
    if kernel in ("adj_weights_e", "adj_weights_h"):
        # lambda (3) + curl coefficient read, six channel weights written, lambda psi read+write
        return cells * 40 + 8 * (pe if kernel == "adj_weights_e" else ph)
    if kernel == "adj_e":
        return cells * 48               # six channel weights read, lambda_H read+write
    if kernel == "adj_h":
        return cells * 52               # six channel weights read, lambda_E read+write, ca
    if kernel == "adj_e_w":
        # gather_e, then the H-side weights: W_E read, lambda_H read+write, hcurl, W_H written
        return cells * 76 + 8 * ph
    if kernel == "material_gradient":
        acc = 8 if args.accumulator == "float64" else 4
        return cells * (2 * comps * hist_b + 12) + ext * 2 * 2 * acc
    if kernel in ("adj_pml_e", "adj_pml_h"):
        return 12 * (pe if kernel == "adj_pml_e" else ph)
    return 0.0


# HBM bytes per launch of a dominant kernel, measured with rocprofv3 in two
# runs (--pmc FETCH_SIZE, then --pmc WRITE_SIZE) and scaled by what the same
# counters read for the 256 MiB device copy bench.py --copy-calibration-mb runs
# first (reads x2.0, writes x1.0). Keyed by (workload, kernel, history dtype,
# accumulator, shots per call): traffic depends on all of them. The commands
# and per-kernel tables are in the profiles/*.md file named with each entry.
MEASURED_TRAFFIC = {
    ("c2", "tm_step_adj", "float32", "float64", 20): (121.9e6, "profiles/c2_fwi_2d_o4_tm_step_kernels.md"),
    ("c2", "tm_step_fwd", "float32", "float64", 20): (90.4e6, "profiles/c2_fwi_2d_o4_tm_step_kernels.md"),
    ("c3", "fwd_e", "float32", "float64", 1): (1656.4e6, "profiles/c3_layered_3d_o4_kernels.md"),
    ("c3", "fwd_h", "float32", "float64", 1): (1383.7e6, "profiles/c3_layered_3d_o4_kernels.md"),
    ("c3", "adj_e_w", "float32", "float64", 1): (2204.0e6, "profiles/c3_layered_3d_o4_kernels.md"),
    ("c3", "adj_h", "float32", "float64", 1): (1520.4e6, "profiles/c3_layered_3d_o4_kernels.md"),
    ("c3", "adj_weights_e", "float32", "float64", 1): (1058.7e6, "profiles/c3_layered_3d_o4_kernels.md"),
}


def _shots_per_call(w, args, free_bytes: int) -> int:
    """Shots per compute() call: all of them when the whole solve fits in 85 %
    of free HBM, else as many as fit (a history larger than HBM runs shot by shot)."""
    from deepgpr_amd.solver.memory import solve_footprint

    shape = [n + 2 * w.pml for n in w.physical_shape] + ([1] if len(w.physical_shape) == 2 else [])
    comps, hist_b, frames = _history_terms(w, args)
    one = solve_footprint(*shape, w.nt, 1, [w.pml] * 6, mode=w.mode, sampling_interval=w.sampling_interval,
                          storage="float32" if args.storage == "int8" else args.storage,
                          compression="int8" if args.storage == "int8" else "none",
                          reconstruct_rhs=frames == 1, fp64_accumulators=args.accumulator == "float64",
                          nrx=w.receiver_location.shape[1])["total"]
    shots = w.source_location.shape[0]
    fit = max(1, int(0.85 * free_bytes // max(one, 1)))
    return min(shots, fit)


# ----------------------------------------------------------------------------- CPU baseline


# CPU sample of each workload: (shots, nt); None = the workload's own value.
# C3's full run needs about 223 GB of host memory per shot for its FP32 E+R
# history and hours of CPU time, so its sample is 40 steps of one shot.
CPU_SAMPLE = {"c1": (1, None), "c2": (1, None), "c3": (1, 40)}


def physical_cores() -> int:
    """Physical cores among the CPUs this process may run on (lscpu's
    CPU,CORE,SOCKET table against sched_getaffinity), 1 thread per core."""
    allowed = os.sched_getaffinity(0)
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        cores = set()
        for line in out.splitlines():
            if line.startswith("#") or not line.strip():
                continue
            cpu, core, sock = (int(x) if x else 0 for x in line.split(",")[:3])
            if cpu in allowed:
                cores.add((sock, core))
        if cores:
            return len(cores)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return len(allowed)


def _cpu_baseline_child(spec: str) -> None:
    """Runs in a subprocess: the reference CPU backend (or the oracle when the
    reference build is absent) on the workload's CPU sample; one warm-up run,
    then the median of three timed runs of forward + backward. Prints one
    JSON line."""
    import statistics

    from deepgpr_amd import synthetic
    from oracle import harness as H

    args = json.loads(spec)
    w = synthetic.WORKLOADS[args["workload"]]() if args["workload"] != "c1" else synthetic.c1_small_2d()
    kind = "reference" if H.reference_available() else "port"
    which = "ref" if kind == "reference" else "oracle"
    shots, nt = CPU_SAMPLE[args["workload"]]
    nt = nt or w.nt
    p = H.workload_problem(w, shots, nt, H.RHS_FROM_E if w.sampling_interval == 1 else 0)
    runs = []
    for r in range(1 + args["repeats"]):
        t = {}
        H.run_cpu(p, which, timing=t)
        if r:
            runs.append(t["seconds"])
    elapsed = statistics.median(runs)
    cells = w.extended_cells()
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    nshot = int(w.source_location.shape[0])
    trunc = f", nt {nt} of {w.nt} (truncated)" if nt < w.nt else f", full nt {nt}"
    print(json.dumps({"value": 2 * cells * nt * shots / elapsed / 1e9, "unit": "Gcell-updates/s",
                      "cores": args["cores"], "threads": threads, "kind": kind,
                      "sample": f"{w.name}: {shots} of {nshot} shots{trunc}, forward+adjoint, FP32 history, "
                                f"OpenMP {threads} threads on {args['cores']} physical cores (close, cores), "
                                f"1 warm-up + median of {len(runs)} runs: "
                                + ", ".join(f"{x:.2f}" for x in runs) + " s"}))


def _cpu_baseline(args):
    cores = physical_cores()
    spec = json.dumps({"workload": args.workload, "repeats": 3, "cores": cores})
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(cores)
    env["OMP_PROC_BIND"] = "close"
    env["OMP_PLACES"] = "cores"
    res = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", spec],
                         capture_output=True, text=True, env=env, timeout=900)
    if res.returncode != 0:
        return {"value": None, "unit": "Gcell-updates/s", "cores": cores, "threads": cores, "kind": None,
                "sample": "failed: " + res.stderr.strip()[-300:]}
    return json.loads(res.stdout.strip().splitlines()[-1])


# ----------------------------------------------------------------------------- main


def main():
    args = _parse()
    if args.cpu_baseline_child is not None:
        _cpu_baseline_child(args.cpu_baseline_child)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported in this build "
              "(shots shard across GPUs with one gradient all-reduce; not run here).", file=sys.stderr)
        return 2
    import torch

    import deepgpr_amd as dg
    from deepgpr_amd.native.loader import get_library

    dev = torch.device("cuda:0")
    if args.copy_calibration_mb:
        n = args.copy_calibration_mb * (1 << 20) // 4
        a = torch.ones(n, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize(dev)
        print(f"[bench] copy calibration: {args.copy_calibration_mb} MiB read + written", file=sys.stderr, flush=True)
        del a, b
    w = _workload(args)
    kw = _compute_kwargs(w, args)
    eps = w.eps_r.to(dev)
    sig = w.sigma.to(dev)
    src_amp = w.source_amplitudes.to(dev)
    src_loc = w.source_location.to(dev)
    rx_loc = w.receiver_location.to(dev)

    nshot = w.source_location.shape[0]
    if not args.shots_per_call:
        args.shots_per_call = _shots_per_call(w, args, torch.cuda.mem_get_info(dev)[0])
    chunk = args.shots_per_call

    starts = list(range(0, nshot, chunk))

    def step(k):
        """Forward + adjoint of batch k mod len(starts); returns its cell-updates."""
        c0 = starts[k % len(starts)]
        e = eps.detach().requires_grad_(True)
        s = sig.detach().requires_grad_(True)
        out = dg.compute(dev, w.dx, w.dt, src_amp, src_loc[c0:c0 + chunk], rx_loc[c0:c0 + chunk], eps_r=e,
                         sigma=s, **kw)
        data = out[-1]
        (data * data).sum().backward()
        del out, data
        return 2 * w.extended_cells() * w.nt * min(chunk, nshot - c0)

    print(f"[bench] {w.name} storage={args.storage} rhs={args.rhs} acc={args.accumulator} "
          f"warmup={args.warmup} steps={args.steps} shots/call={chunk}", file=sys.stderr, flush=True)
    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize(dev)
    updates = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        updates += step(args.warmup + k)
        torch.cuda.synchronize(dev)
        print(f"[bench] step {k + 1}/{args.steps} done", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / max(args.steps, 1)
    value = updates / max(t1 - t0, 1e-9) / 1e9

    roofline = None
    kernels = {}
    if not args.no_kernel_timing:
        lib = get_library()
        lib.dgx_set_kernel_timing(1)
        step(0)  # per-launch averages: one compute() call is enough
        torch.cuda.synchronize(dev)
        lib.dgx_set_kernel_timing(0)
        for kid in range(64):
            name = lib.dgx_kernel_name(kid)
            if not name:
                break
            tot, cnt = ctypes.c_double(), ctypes.c_longlong()
            lib.dgx_kernel_timing(kid, ctypes.byref(tot), ctypes.byref(cnt))
            if cnt.value:
                kernels[name.decode()] = (tot.value, cnt.value)
        lib.dgx_set_kernel_timing(0)
        ranked = sorted(((t, n) for n, (t, c) in kernels.items() if kernel_bytes(n, w, args) > 0), reverse=True)
        if ranked:
            _, dom = ranked[0]
            tot, cnt = kernels[dom]
            avg_s = tot / cnt / 1e3
            bytes_launch = kernel_bytes(dom, w, args)
            achieved = bytes_launch / avg_s / 1e9
            measured = MEASURED_TRAFFIC.get((args.workload, dom, args.storage, args.accumulator, chunk))
            roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": int(measured[0]) if measured else None,
                        "traffic_kind": "profile constant (rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this kernel "
                                        "and configuration, not of this run)" if measured else None,
                        "avg_launch_us": round(avg_s * 1e6, 2), "algorithmic_bytes_per_launch": int(bytes_launch)}
            if measured:
                roofline["traffic_source"] = measured[1]
    fwd_b, adj_b = model_bytes_per_cell(w, args)
    pair_bytes = fwd_b + adj_b
    cpu = None if args.no_cpu_baseline else _cpu_baseline(args)
    line = {
        "metric": "Gcell-updates/s fwd+adjoint (% HBM roofline); serve p50/p99 ms vs plain solve",
        "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "shots": int(w.source_location.shape[0]), "shots_per_step": chunk, "nt": w.nt,
                   "fdtd_order": w.fdtd_order, "history": args.storage, "rhs_history": args.rhs,
                   "gradient_accumulator": args.accumulator, "sampling_interval": w.sampling_interval},
        "roofline": roofline,
        "step_roofline": {"bytes_per_cell_update": [round(fwd_b, 2), round(adj_b, 2)],
                          "roofline_gcell_per_s": round(2 * HBM_PEAK_GBS / pair_bytes, 1),
                          "frac": round(value * 1e9 / 2 * pair_bytes / (HBM_PEAK_GBS * 1e9), 4)},
        "kernels_ms_per_launch": {n: round(t / c, 4) for n, (t, c) in kernels.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
