"""Gradient accuracy of the history formats against the oracle run in
float64 (float64 fields, a lossless float64 history and float64 sums), and a
finite-difference check of the adjoint on the GPU.

Targets: BF16 histories within 1e-2 rel-L2 (cosine >= 0.999) of the float64
gradient at <= 55 % of the FP32-history memory (the BF16 E-only history,
50 %, against the FP32 E-only history compute() keeps by default for C2);
FP32 histories within the FP32 tolerance (1e-4 rel-L2 for gradients). Each
case prints one JSON line with its figures.
"""
import json

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _metrics(x, ref):
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / den), float(x @ ref / (np.linalg.norm(x) * den))


def _history_bytes(p):
    frames = p.nt_saved
    return p.history_bytes(frames, 4) + (p.history_bytes(1, 4) if p.e_only() else p.history_bytes(frames, 4))


def _formats(h, p, formats, ref, label):
    out = {}
    for name, st in formats:
        q = h.with_storage(p, storage_type=st)
        gpu = h.run_gpu(q, api="dgx", fp64=True)
        row = {"case": label, "history": name, "memory": _history_bytes(q)}
        for key in ("g_eps", "g_sig"):
            row[key] = _metrics(gpu[key], ref[key])
        row["rx"] = _metrics(gpu["rx"], ref["rx"])[0]
        print(json.dumps(row), flush=True)
        out[name] = row
    return out


def test_c2_bf16_history_accuracy_and_memory(oracle_built, cuda_device):
    """C2 geometry, the first 4 shots, nt 300, FP64 accumulators."""
    from deepgpr_amd import synthetic
    h = oracle_built
    print("[gpu accuracy] C2 histories", flush=True)
    p = h.workload_problem(synthetic.c2_fwi_2d(), 4, 300, h.RHS_FROM_E)
    ref = h.run_cpu(p, "oracle", real=np.float64)
    rows = _formats(h, p, [("float32 E-only", h.RHS_FROM_E), ("bfloat16 E-only", h.KIND_BF16 | h.RHS_FROM_E),
                           ("bfloat16 E+R", h.KIND_BF16), ("int8 8x8 E+R", h.int8_storage_type((8, 8, 1)))],
                    ref, "c2 4 shots nt 300")
    f32, bf16 = rows["float32 E-only"], rows["bfloat16 E-only"]
    for key in ("g_eps", "g_sig"):
        assert f32[key][0] <= 1e-5, (key, f32[key])
        assert bf16[key][0] <= 1e-2 and bf16[key][1] >= 0.999, (key, bf16[key])
    assert bf16["memory"] <= 0.55 * f32["memory"]


def test_c3_physics_history_formats(oracle_built, cuda_device):
    """C3's layered-plus-spheres model, order 4, CPML 16, S 4, on a 96^3 grid
    (128^3 cells with the CPML), nt 400: FP32, BF16 and INT8 (4x4x4 tiles)
    E+R histories against the float64 oracle."""
    from deepgpr_amd import synthetic
    h = oracle_built
    print("[gpu accuracy] C3 physics histories", flush=True)
    w = synthetic.c3_layered_3d(nshot=1, nt=400, shape=(96, 96, 96))
    p = h.workload_problem(w, 1, 400, 0)
    assert (p.NX, p.S, p.order) == (129, 4, 4)
    ref = h.run_cpu(p, "oracle", real=np.float64)
    rows = _formats(h, p, [("float32 E+R", 0), ("bfloat16 E+R", h.KIND_BF16),
                           ("int8 4x4x4 E+R", h.int8_storage_type((4, 4, 4)))], ref, "c3 physics 96^3 nt 400")
    for key in ("g_eps", "g_sig"):
        # FP32 fields alone (receivers ~4e-6 from the float64 oracle) bound the FP32 history's gradients
        assert rows["float32 E+R"][key][0] <= 1e-4, key
        assert rows["bfloat16 E+R"][key][0] <= 1e-2 and rows["bfloat16 E+R"][key][1] >= 0.999, key
        assert rows["int8 4x4x4 E+R"][key][0] <= 5e-2 and rows["int8 4x4x4 E+R"][key][1] >= 0.99, key


def test_c2_finite_differences_on_16_random_cells(cuda_device):
    """Central differences of a data misfit (evaluated in float64 from the
    GPU receivers) against the adjoint eps_r and sigma gradients at 16 random
    cells of the C2 model (seed 2029) that the first shot's wave reaches in
    600 steps. As in the reference's Taylor checks (tests/test_discrete_adjoint.py:
    204-282, 510-521) the best of a few step sizes must agree to 1.5e-2."""
    import deepgpr_amd as dg
    from deepgpr_amd import synthetic
    print("[gpu accuracy] finite differences on 16 C2 cells", flush=True)
    w = synthetic.c2_fwi_2d()
    nt = 600
    dev = cuda_device
    kw = dict(dx=w.dx, dt=w.dt, source_amplitudes=w.source_amplitudes[:, :nt].to(dev),
              source_location=w.source_location[:1].to(dev), receiver_location=w.receiver_location[:1].to(dev),
              pmlthick=w.pml, fdtd_order=w.fdtd_order, mode=w.mode, gradient_accumulator="float64")
    eps0, sig0 = w.eps_r.to(dev), w.sigma.to(dev)
    with torch.no_grad():
        base = dg.compute(dev, eps_r=eps0, sigma=sig0, **kw)[-1].double()
    scale = base.abs().max()
    target = 0.85 * base

    def misfit(eps, sig):
        rx = dg.compute(dev, eps_r=eps, sigma=sig, **kw)[-1].double()
        return 0.5 * ((rx - target) / scale).square().sum()

    e = eps0.clone().requires_grad_(True)
    s = sig0.clone().requires_grad_(True)
    misfit(e, s).backward()
    gen = torch.Generator().manual_seed(2029)
    # in 600 steps the misfit is sensitive to the cells within ~16 rows of the
    # surface and ~64 columns of the source (physical cell (5, 10)) only
    cand = torch.stack([torch.randint(1, 17, (256,), generator=gen), torch.randint(0, 64, (256,), generator=gen)], 1)
    src = tuple(int(v) for v in w.source_location[0, 0, :2])

    def pick(grad):
        """The first 16 distinct candidates (in their seeded order), the
        source cell aside, whose perturbed values stay valid for every step
        (eps_r >= 1, sigma >= 0) and whose gradient is at least 1 % of the
        largest, so that both sides are well above round-off."""
        floor = 1e-2 * float(grad.abs().max())
        out = []
        for x, y in cand.tolist():
            if ((x, y) not in out and (x, y) != src and float(eps0[x, y]) > 1.5 and float(sig0[x, y]) > 2e-3
                    and abs(float(grad[x, y])) >= floor):
                out.append((x, y))
        assert len(out) >= 16, out
        return out[:16]

    def check(x, y, which):
        grad, steps = (e.grad, (0.2, 0.1, 0.05)) if which == "eps" else (s.grad, (1e-3, 5e-4, 2.5e-4))
        adj = float(grad[x, y])
        best, best_fd = float("inf"), None
        for hstep in steps:
            with torch.no_grad():
                delta = torch.zeros_like(eps0)
                delta[x, y] = hstep
                if which == "eps":
                    lp, lm = float(misfit(eps0 + delta, sig0)), float(misfit(eps0 - delta, sig0))
                else:
                    lp, lm = float(misfit(eps0, sig0 + delta)), float(misfit(eps0, sig0 - delta))
            fd = (lp - lm) / (2 * hstep)
            err = abs(fd - adj) / max(abs(fd), abs(adj), 1e-30)
            if err < best:
                best, best_fd = err, fd
        return {"cell": [x, y], "adjoint": adj, "fd": best_fd, "rel_err": best}

    rows = {which: [check(x, y, which) for x, y in pick(e.grad if which == "eps" else s.grad)]
            for which in ("eps", "sigma")}
    print(json.dumps({"case": "c2 finite differences, 16 cells each", **rows}), flush=True)
    for which in ("eps", "sigma"):
        assert max(r["rel_err"] for r in rows[which]) < 1.5e-2, rows[which]
