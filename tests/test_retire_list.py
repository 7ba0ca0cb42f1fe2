"""The batch runtime's retire list (deepgpr_amd/csrc/dgx_retire.h), driven on
the host with fake completion markers: retired memory is released only after
its marker completes, and a marker that cannot be recorded drains the stream
(or keeps the memory) before any release. No GPU needed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_retire_list_with_fake_events(tmp_path):
    exe = tmp_path / "retire_list_check"
    src = os.path.join(ROOT, "tests", "native", "retire_list_check.cpp")
    inc = os.path.join(ROOT, "deepgpr_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-fsanitize=address,undefined", "-I" + inc,
                    src, "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all checks passed" in res.stdout
