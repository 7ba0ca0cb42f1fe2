"""Build the in-tree HIP library for gfx950: ``python -m deepgpr_amd.build``."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ("csrc/dgx_kernels.hip", "csrc/dgx_xmarch.hip", "csrc/dgx_tm2d.hip", "csrc/dgx_tm2d_fused.hip",
           "csrc/dgx_api.hip")
HIPCC_FLAGS = (
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # keep a*b+c unfused so the FP32 kernels reproduce the reference CPU bitwise
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
)


def build(verbose: bool = True) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = os.path.join(HERE, "libdeepgpr_amd.so")
    tmp = out + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I" + os.path.join(ROOT, "include"), "-o", tmp,
           *(os.path.join(HERE, s) for s in SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errors = "\n".join(line for line in res.stderr.splitlines() if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errors}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build())
    sys.exit(0)
