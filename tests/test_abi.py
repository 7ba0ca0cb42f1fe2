"""The C ABI: header, ctypes tables and the built library agree
(restates the reference's tests/test_native_abi.py:58-136 for this build)."""
import ctypes
import os
import re
import subprocess

import pytest

from deepgpr_amd.native import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "deepgpr_amd.h")


def _declarations():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"DGX_API\s+([\w\s\*]+?)\b(\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        name, params = m.group(2), m.group(3).strip()
        if params in ("", "void"):
            decls[name] = []
        else:
            decls[name] = [re.split(r"[\s\*]+", p.strip())[-1] for p in params.split(",")]
    return decls


def test_header_matches_ctypes_tables():
    decls = _declarations()
    assert set(decls) == set(abi.NATIVE_FUNCTIONS)
    for name, (params, _ret) in abi.NATIVE_FUNCTIONS.items():
        if name in ("forward", "backward"):
            got = decls[name]
            want = [p for p, _ in params]
            # the header spells the backward lambda phi names in full
            assert len(got) == len(want), name
            assert [g.replace("lambda_", "") for g in got] == [w.replace("lambda_", "") for w in want], name
        else:
            assert len(decls[name]) == len(params), name


def test_abi6_argument_counts():
    assert len(abi.FORWARD_SIGNATURE) == 81
    assert len(abi.BACKWARD_SIGNATURE) == 86


def test_struct_layout_matches_header_order():
    text = open(HEADER).read()
    body = text[text.index("typedef struct dgx_solve {"):text.index("} dgx_solve;")]
    names = []
    for line in body.splitlines()[1:]:
        line = re.sub(r"/\*.*?\*/", "", line).strip()
        if not line:
            continue
        for part in line.rstrip(";").split(","):
            token = part.strip().split()[-1].lstrip("*")
            names.append(token.split("[")[0])
    assert names == [f for f, _ in abi.DgxSolve._fields_]


def test_library_loads_and_exports_every_declared_symbol():
    from deepgpr_amd.native.loader import get_library, library_path
    if not os.path.isfile(library_path()):
        pytest.skip("HIP library not built")
    lib = get_library()
    nm = subprocess.run(["nm", "-D", "--defined-only", library_path()], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in nm.splitlines() if line.strip()}
    for name in _declarations():
        assert name in exported, name
    assert lib.deepgpr_abi_version() == 6
    assert lib.dgx_abi_version() == abi.DGX_ABI_VERSION
    for probe in ("deepgpr_supports_error_reporting", "deepgpr_deterministic_adjoint", "deepgpr_supports_external_pml",
                  "deepgpr_supports_pml_material_gradients", "deepgpr_supports_int8_wavefield",
                  "deepgpr_supports_rhs_reconstruction", "deepgpr_supports_physical_history",
                  "deepgpr_supports_tm2d_fast_path"):
        assert getattr(lib, probe)() == 1


def test_error_path_without_gpu_work():
    """A malformed descriptor fails on the host with a message and launches nothing."""
    from deepgpr_amd.native.loader import get_library, library_path
    if not os.path.isfile(library_path()):
        pytest.skip("HIP library not built")
    lib = get_library()
    s = abi.DgxSolve()
    s.abi_version = abi.DGX_ABI_VERSION + 7
    assert lib.dgx_forward(ctypes.byref(s), None) == -1
    assert b"abi_version" in lib.dgx_last_error()
    s.abi_version = abi.DGX_ABI_VERSION
    s.nx_fields = s.ny_fields = s.nz_fields = 1
    assert lib.dgx_backward(ctypes.byref(s), None) == -1
    assert b"field sizes" in lib.dgx_last_error()
    lib.deepgpr_clear_last_error()
    assert lib.deepgpr_last_error() == b""


def test_cpu_device_is_refused():
    import torch
    import deepgpr_amd as dg
    with pytest.raises(dg.NativeLibraryError):
        dg.compute("cpu", dx=0.02, dt=3e-11, source_amplitudes=torch.zeros(1, 10, 1),
                   source_location=torch.tensor([[[2, 2, 0]]]), receiver_location=torch.tensor([[[3, 3, 0]]]),
                   eps_r=torch.ones(8, 8), sigma=torch.zeros(8, 8), pmlthick=2)


REF_HEADER = "/root/reference/src/DeepGPR/lib/deepgpr.h"


def _prototypes(path, macro):
    """name -> [(kind, parameter name)] of every `macro` declaration; kind is
    the parameter's type with spaces normalised (const and pointers kept)."""
    text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
    out = {}
    for m in re.finditer(macro + r"\s+([\w\s\*]+?)\b(\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        params = m.group(3).strip()
        plist = []
        if params not in ("", "void"):
            for p in params.split(","):
                p = " ".join(p.replace("*", " * ").split())
                *kind, name = p.split(" ")
                plist.append((" ".join(kind).replace(" *", "*"), name))
        out[m.group(2)] = (" ".join(m.group(1).replace("*", " * ").split()).replace(" *", "*"), plist)
    return out


@pytest.mark.skipif(not os.path.isfile(REF_HEADER), reason="reference sources not present")
def test_abi6_exports_match_the_reference_header():
    """Every export of the reference deepgpr.h (deepgpr.h:48-152) is declared
    in include/deepgpr_amd.h with the same return kind and the same parameter
    kinds, names and order, so the library binds as a drop-in deepgpr.so."""
    ref = _prototypes(REF_HEADER, "DEEPGPR_API")
    ours = _prototypes(HEADER, "DGX_API")
    assert len(ref) >= 17
    for name, (ret, params) in ref.items():
        assert name in ours, name
        oret, oparams = ours[name]
        assert oret == ret, name
        assert [k for k, _ in oparams] == [k for k, _ in params], name
        assert [n for _, n in oparams] == [n for _, n in params], name
    assert len(ref["forward"][1]) == 81 and len(ref["backward"][1]) == 86
