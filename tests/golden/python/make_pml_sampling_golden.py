"""Writes tests/golden/python/pml_sampling_golden.npz: input/output vectors of two
Python functions of the reference that no reference test stores values for,

- build_pml_coeffs (src/DeepGPR/solver/pml.py:227-298): the float32/float64
  CPML descriptors and update coefficients, whose bits depend on torch's
  linspace/pow sequence;
- source_max_frequency and recommended_sampling_interval
  (src/DeepGPR/solver/sampling.py:58-134): the "auto" sampling interval.

Run once, in a container holding the reference checkout, with
``python tests/golden/python/make_pml_sampling_golden.py [reference_src_dir]``; the
test suite only reads the .npz. The reference modules are loaded one by one
under empty parent packages, so no package ``__init__`` runs and no native
library of the reference is loaded.
"""
from __future__ import annotations

import importlib
import math
import os
import sys
import types

import numpy as np
import torch

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pml_sampling_golden.npz")

# (extended nx, ny, nz, six thicknesses, dx, dt, max_velocity or None, dtype)
PML_CASES = [
    (60, 50, 40, (10, 10, 10, 10, 10, 10), 0.02, 3.0e-11, None, "float32"),
    (23, 26, 1, (3, 5, 4, 2, 0, 0), (0.020, 0.016, 0.012), 2.5e-11, None, "float32"),
    (288, 288, 224, (16, 16, 16, 16, 16, 16), 0.02, 2.5e-11, None, "float32"),
    (440, 240, 1, (20, 20, 20, 20, 0, 0), 0.02, 3.0e-11, None, "float32"),
    (30, 30, 30, (2, 1, 3, 2, 1, 2), (0.020, 0.018, 0.016), 1.5e-11, 1.5e8, "float32"),
    (40, 40, 1, (8, 6, 7, 5, 0, 0), 0.01, 1.5e-11, None, "float64"),
]

# (kind, length, dt, parameters): waveforms built here in float64, cast to float32
SAMPLING_CASES = [
    ("ricker", 2000, 3.0e-11, (2.0e8, 5e-9)),
    ("ricker", 3000, 3.0e-11, (2.5e8, 4e-9)),
    ("ricker", 2000, 2.5e-11, (3.0e8, 1.0 / 3.0e8)),
    ("ricker", 800, 1.2e-11, (8.0e8, 1.5e-9)),
    ("gauss_sine", 1500, 2e-11, (4.0e8, 2e-9)),
    ("two_rows", 1200, 3.0e-11, (1.0e8, 4.0e8)),
    ("zero", 500, 3.0e-11, ()),
]


def waveform(kind: str, n: int, dt: float, par) -> torch.Tensor:
    t = torch.arange(n, dtype=torch.float64) * dt
    if kind == "ricker":
        f, peak = par
        a = (math.pi * f * (t - peak)) ** 2
        return ((1 - 2 * a) * torch.exp(-a)).float().reshape(1, n, 1)
    if kind == "gauss_sine":
        f, peak = par
        return (torch.sin(2 * math.pi * f * t) * torch.exp(-((t - peak) * f) ** 2)).float().reshape(1, n, 1)
    if kind == "two_rows":
        rows = []
        for f in par:
            a = (math.pi * f * (t - 1.2 / f)) ** 2
            rows.append((1 - 2 * a) * torch.exp(-a))
        return torch.stack(rows).float().reshape(2, n, 1)
    return torch.zeros(1, n, 1)


def _load_reference(src: str):
    root = os.path.join(src, "DeepGPR")
    for name, sub in (("DeepGPR", ""), ("DeepGPR.config", "config"), ("DeepGPR.preprocessing", "preprocessing"),
                      ("DeepGPR.utils", "utils"), ("DeepGPR.solver", "solver")):
        pkg = types.ModuleType(name)
        pkg.__path__ = [os.path.join(root, sub)]
        sys.modules[name] = pkg
    return importlib.import_module("DeepGPR.solver.pml"), importlib.import_module("DeepGPR.solver.sampling")


def main() -> int:
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src"
    pml, sampling = _load_reference(src)
    out = {}
    for k, (nx, ny, nz, th, dx, dt, vmax, dt_name) in enumerate(PML_CASES):
        dtype = getattr(torch, dt_name)
        kw = {} if vmax is None else {"max_velocity": vmax}
        res = pml.build_pml_coeffs(None, None, dt, dx, nx, ny, nz, torch.tensor(th), torch.device("cpu"), dtype, **kw)
        for i, t in enumerate(res):
            out[f"pml{k}_{i}"] = t.detach().cpu().numpy()
    for k, (kind, n, dt, par) in enumerate(SAMPLING_CASES):
        w = waveform(kind, n, dt, par)
        out[f"wave{k}"] = w.numpy()
        out[f"fmax{k}"] = np.float64(sampling.source_max_frequency(w, dt))
        out[f"interval{k}"] = np.int64(sampling.recommended_sampling_interval(w, dt))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(out)} arrays")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
