"""The block -> cell mapping of the 3D sweeps (deepgpr_amd/csrc/dgx_sweep.h),
checked on the host: every cell advanced exactly once, the inverse used for
the source lists, the interior-box conditions and the multiply-high
divisions. No GPU needed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_sweep_map_covers_every_cell_once(tmp_path):
    exe = tmp_path / "sweep_map_check"
    src = os.path.join(ROOT, "tests", "native", "sweep_map_check.cpp")
    inc = os.path.join(ROOT, "deepgpr_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-fsanitize=address,undefined", "-I" + inc,
                    src, "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    assert "all checks passed" in res.stdout
