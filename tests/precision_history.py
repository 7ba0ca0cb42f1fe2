"""Gradient error of each wavefield-history format against the float64 oracle.

Runs the first ``--shots`` shots of the 2D multi-shot FWI workload
(400x200 cells, order 4, CPML 20, mode 3, nt 3000) through the HIP library's
C ABI with FP32, BF16, FP16 and INT8 histories, and runs the oracle in float64
(FP64 fields and a lossless FP64 history) on the same native inputs. Prints
one JSON line per history format:

- the relative L2 error and the cosine of grad_eps, grad_sigma and the
  receiver traces against the float64 oracle;
- the history bytes relative to the FP32 E-only history compute() uses by
  default for this workload.

The adjoint source is 2 x the float64 oracle's receiver traces (the
gradient of sum(receivers^2)), identical for every run. The oracle part needs
no GPU and about 8 GB of host memory per shot; it can run separately:

    python tests/precision_history.py --shots 4 --oracle-out ORACLE.npz   # CPU
    python tests/precision_history.py --oracle-in ORACLE.npz              # GPU
    python tests/precision_history.py --shots 4                           # both
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from deepgpr_amd import synthetic  # noqa: E402
from oracle import harness as H  # noqa: E402


def _metrics(x: np.ndarray, ref: np.ndarray) -> dict:
    x = x.astype(np.float64).ravel()
    ref = ref.astype(np.float64).ravel()
    den = np.linalg.norm(ref)
    cos = float(x @ ref / max(np.linalg.norm(x) * den, 1e-300))
    return {"rel_l2": float(np.linalg.norm(x - ref) / max(den, 1e-300)), "cosine": cos}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--shots", type=int, default=4)
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--workload", default="c2", choices=("c1", "c2"))
    ap.add_argument("--threads", type=int, default=16, help="OpenMP threads of the float64 oracle")
    ap.add_argument("--oracle-out", default=None, help="run only the float64 oracle and save its results here")
    ap.add_argument("--oracle-in", default=None, help="load the float64 oracle results instead of running it")
    args = ap.parse_args()

    w = synthetic.WORKLOADS[args.workload]() if args.workload != "c1" else synthetic.c1_small_2d()
    if w.sampling_interval != 1:
        raise SystemExit("the E-only baseline needs sampling interval 1")
    if args.oracle_in:
        with np.load(args.oracle_in, allow_pickle=False) as z:
            saved = {k: z[k] for k in z.files}
        shots, nt = int(saved["shots"]), int(saved["nt"])
        if saved["workload"].item() != w.name:
            raise SystemExit(f"{args.oracle_in} holds {saved['workload'].item()}, not {w.name}")
    else:
        shots, nt = min(args.shots, w.source_location.shape[0]), args.nt
    base = H.workload_problem(w, shots, nt, H.RHS_FROM_E)

    if args.oracle_in:
        base.data_grad = saved["data_grad"]
        ref_rx, ref_ge, ref_gs = saved["rx"], saved["g_eps"], saved["g_sig"]
    else:
        # adjoint source: gradient of sum(rx^2) at the float64 traces
        t0 = time.perf_counter()
        fwd = H.run_cpu(H.with_storage(base, storage_type=H.NO_HISTORY), "oracle", real=np.float64,
                        backward=False, threads=args.threads)
        base.data_grad = (2.0 * fwd["rx"]).astype(np.float32)
        del fwd
        ref = H.run_cpu(H.with_storage(base, storage_type=H.RHS_FROM_E), "oracle", real=np.float64,
                        threads=args.threads)
        print(f"[precision] {w.name}: {shots} shots, nt={nt}: float64 oracle {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        ref_rx, ref_ge, ref_gs = ref["rx"], ref["g_eps"], ref["g_sig"]
        del ref
        if args.oracle_out:
            np.savez_compressed(args.oracle_out, workload=np.array(w.name), shots=shots, nt=nt,
                                data_grad=base.data_grad, rx=ref_rx, g_eps=ref_ge, g_sig=ref_gs)
            return 0

    fp32_bytes = base.history_bytes(base.nt_saved, 4) + base.history_bytes(1, 4)
    formats = [
        ("float32 E-only, FP64 acc", H.RHS_FROM_E, True),
        ("float32 E-only, FP32 acc", H.RHS_FROM_E, False),
        ("float16 E-only, FP64 acc", H.KIND_F16 | H.RHS_FROM_E, True),
        ("bfloat16 E-only, FP64 acc", H.KIND_BF16 | H.RHS_FROM_E, True),
        ("bfloat16 E+R, FP64 acc", H.KIND_BF16, True),
        ("int8 8x8 E+R, FP64 acc", H.int8_storage_type((8, 8, 1)), True),
    ]
    for name, st, fp64 in formats:
        p = H.with_storage(base, storage_type=st)
        frames = p.nt_saved
        hist = p.history_bytes(frames, 4) + (p.history_bytes(1, 4) if p.e_only() else p.history_bytes(frames, 4))
        t0 = time.perf_counter()
        out = H.run_gpu(p, backward=True, api="dgx", fp64=fp64)
        line = {"format": name, "workload": w.name, "shots": shots, "nt": nt, "history_bytes": int(hist),
                "history_vs_fp32_e_only": round(hist / fp32_bytes, 4),
                "grad_eps": _metrics(out["g_eps"], ref_ge), "grad_sigma": _metrics(out["g_sig"], ref_gs),
                "receivers": _metrics(out["rx"], ref_rx), "gpu_s": round(time.perf_counter() - t0, 2)}
        print(json.dumps(line), flush=True)
        del out
    return 0


if __name__ == "__main__":
    sys.exit(main())
