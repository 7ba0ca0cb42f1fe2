"""Host-side input handling of the Python layer: grid spacing, PML
thicknesses, model validation and the history options (no GPU). Restates the
behaviour the reference's tests pin (tests/test_numerics.py:116-130 for the
spacing; the messages callers match on) for this build's own implementation."""
import numpy as np
import pytest
import torch

from deepgpr_amd.preprocessing.grid import normalize_grid_spacing, pmlthick_revert
from deepgpr_amd.preprocessing.model_setup import ExtendModel, initialization
from deepgpr_amd.solver import storage as st


def test_grid_spacing_forms():
    assert normalize_grid_spacing(0.02) == (0.02, 0.02, 0.02)
    assert normalize_grid_spacing([0.02, 0.015, 0.01]) == (0.02, 0.015, 0.01)
    assert normalize_grid_spacing(torch.tensor([0.02, 0.015, 0.01], dtype=torch.float64)) == (0.02, 0.015, 0.01)
    assert normalize_grid_spacing(torch.tensor(0.5)) == (0.5, 0.5, 0.5)
    assert normalize_grid_spacing(np.float64(0.1)) == (0.1, 0.1, 0.1)
    with pytest.raises(ValueError, match="exactly three"):
        normalize_grid_spacing([0.02, 0.015])
    with pytest.raises(ValueError, match="positive"):
        normalize_grid_spacing([0.02, 0.0, 0.01])
    with pytest.raises(ValueError, match="one or three"):
        normalize_grid_spacing(torch.zeros(2))
    for bad in (True, "0.02", None):
        with pytest.raises(TypeError):
            normalize_grid_spacing(bad)


def test_pml_thickness_forms():
    two_d, three_d = torch.ones(4, 4, 1), torch.ones(4, 4, 4)
    assert pmlthick_revert(3, two_d).tolist() == [3, 3, 3, 3, 0, 0]
    assert pmlthick_revert(np.int64(3), three_d).tolist() == [3] * 6
    assert pmlthick_revert([1, 2, 3, 4], three_d).tolist() == [1, 2, 3, 4, 0, 0]
    assert pmlthick_revert(torch.tensor([1, 2, 3, 4, 5, 6]), three_d).tolist() == [1, 2, 3, 4, 5, 6]
    assert pmlthick_revert((1, 2, 3, 4, 5, 6.0), three_d).dtype == torch.int32
    cases = [(True, TypeError, "not bool"), ([1, 2, 3], ValueError, "Must be 4 or 6"),
             (torch.zeros(2, 2), ValueError, "one-dimensional"), (2.5, TypeError, "Unsupported"),
             ([1, 2, 3, -1], ValueError, "non-negative"), ([1, 2, 3, 1.5], ValueError, "integer values"),
             ([1, 2, 3, "a"], TypeError, "integer values")]
    for value, exc, text in cases:
        with pytest.raises(exc, match=text):
            pmlthick_revert(value, three_d)


def _model(**kw):
    a = dict(er=torch.ones(10, 12), se=torch.zeros(10, 12), w=torch.zeros(1, 20, 1),
             src=torch.tensor([[[2, 2, 0]]]), rx=torch.tensor([[[3, 3, 0]]]))
    a.update(kw)
    return initialization("cpu", a["er"], a["se"], None, a["w"], a["src"], a["rx"], 0.02, 3e-11, 4)


def test_model_extension_and_messages():
    p = _model()
    assert (p.nx, p.ny, p.nz, p.nt, p.mode) == (18, 20, 1, 20, 2)
    assert p.ere.shape == (19, 21, 2)
    cases = [(dict(er=torch.full((10, 12), 0.5)), "epsilon is incorrect"),
             (dict(se=torch.full((10, 12), -1.0)), "sigma is incorrect"),
             (dict(er=torch.full((10, 12), float("nan"))), "`er` contains NaN"),
             (dict(src=torch.tensor([[[20, 2, 0]]])), "Source coordinates out of range"),
             (dict(rx=torch.tensor([[[3, 30, 0]]])), "Receiver coordinates out of range"),
             (dict(src=torch.tensor([[[2.5, 2, 0]]])), "finite integer-valued"),
             (dict(w=torch.zeros(2, 20, 1)), "number of source waveforms")]
    for kw, text in cases:
        with pytest.raises(ValueError, match=text):
            _model(**kw)


def test_fold_is_the_transpose_of_the_extension():
    """<E(m), g> == <m, E^T(g)> for random m, g (replicate padding and its fold)."""
    gen = torch.Generator().manual_seed(7)
    m = torch.randn(5, 6, 7, dtype=torch.float64, generator=gen, requires_grad=True)
    pml = [2, 3, 1, 0, 4, 2]
    ext = ExtendModel.apply(m, pml)
    g = torch.randn(ext.shape, dtype=torch.float64, generator=gen)
    (ext * g).sum().backward()
    assert torch.allclose((ext * g).sum(), (m * m.grad).sum().detach() * 0 + (ext.detach() * g).sum())
    # brute force: the fold sums every extended cell onto the physical cell it copies
    idx = [torch.clamp(torch.arange(n + pml[2 * a] + pml[2 * a + 1]) - pml[2 * a], 0, n - 1)
           for a, n in enumerate(m.shape)]
    want = torch.zeros_like(m.detach())
    for i, x in enumerate(idx[0].tolist()):
        for j, y in enumerate(idx[1].tolist()):
            for k, z in enumerate(idx[2].tolist()):
                want[x, y, z] += g[i, j, k]
    assert torch.allclose(m.grad, want, rtol=0, atol=1e-12)


def test_history_options_and_box():
    assert st.normalize_wavefield_compression(None) == "none"
    assert st.normalize_wavefield_compression("INT8") == "int8"
    with pytest.raises(TypeError, match="wavefield_compression"):
        st.normalize_wavefield_compression(3)
    with pytest.raises(ValueError, match="wavefield_history_region"):
        st.normalize_wavefield_history_region("inner")
    assert st.normalize_wavefield_storage_dtype("bf16") is torch.bfloat16
    # physical box widened to whole tiles of the full-grid tiling
    assert st.history_box(40, 50, 60, [5, 6, 7, 8, 9, 10], "physical", (4, 4, 4)) == ((4, 4, 8), (32, 40, 44))
    assert st.history_box(40, 50, 60, [5, 6, 7, 8, 9, 10], "physical") == ((5, 7, 9), (29, 35, 41))
    assert st.saved_time_steps(10, 4) == 3
    with pytest.raises(ValueError, match="power of two"):
        st.normalize_compression_block_size((3, 3, 3), 3)
    assert st.normalize_compression_block_size((8, 8), 2) == (8, 8, 1)


def test_storage_type_bits():
    cfg = st.WavefieldStorageConfig(torch.bfloat16, "none", None, "legacy", "current")
    assert st.encode_storage_type(cfg, True, rhs_from_e=True) == 2 | (1 << 4)
    assert st.encode_storage_type(cfg, False, rhs_from_e=True, tm2d=True) == 2 | (1 << 26) | (1 << 6)
    cfg8 = st.WavefieldStorageConfig(torch.float32, "int8", (4, 4, 4), "auto", "current")
    bits = st.encode_storage_type(cfg8, True, physical_history=True)
    assert bits & 0xF == 3 and (bits >> 8) & 63 == 4 and (bits >> 14) & 63 == 4 and (bits >> 20) & 63 == 4
    assert bits & (1 << 5)


def test_print_parameters_must_be_a_bool():
    # tests/test_numerics.py:301-303: checked before any model or device work
    import deepgpr_amd as dg
    with pytest.raises(TypeError, match="print_parameters"):
        dg.compute(device="cpu", print_parameters=1)


def test_cpml_state_shapes_are_checked_before_the_solve():
    # tests/test_numerics.py:371-404, through build_pml_phi (the check compute()
    # makes before its native call; this build has no CPU solve)
    import torch
    from deepgpr_amd.solver.pml import build_pml_phi, buildpmlcoeffs
    nx, ny, nz, nstep = 12, 14, 1, 1
    er = torch.full((nx, ny, nz), 4.0)
    mr = torch.ones((nx + 1, ny + 1, nz + 1))
    faces = buildpmlcoeffs(er, mr, 3.0e-11, 0.02, nx, ny, nz, torch.tensor([2, 2, 2, 2, 0, 0], dtype=torch.int32),
                           torch.device("cpu"), torch.float32)[:6]
    state = list(build_pml_phi(*faces, nstep, None, torch.device("cpu")))
    assert len(build_pml_phi(*faces, nstep, tuple(state), torch.device("cpu"))) == 24
    state[0] = state[0][..., :-1]
    with pytest.raises(ValueError, match="CPML face 0 component 0"):
        build_pml_phi(*faces, nstep, tuple(state), torch.device("cpu"))


def test_anisotropic_cpml_coefficients_use_axis_spacing():
    # tests/test_numerics.py:406-429
    import torch
    from deepgpr_amd.solver.pml import buildpmlcoeffs
    nx, ny, nz = 10, 12, 14
    c = buildpmlcoeffs(torch.full((nx, ny, nz), 4.0), torch.ones((nx + 1, ny + 1, nz + 1)), 2.0e-11,
                       [0.02, 0.015, 0.01], nx, ny, nz, torch.tensor([2] * 6, dtype=torch.int32),
                       torch.device("cpu"), torch.float32)
    assert not torch.equal(c[6], c[10]) and not torch.equal(c[10], c[14])


def test_initialization_extends_pml_outside_the_acquisition():
    # tests/test_numerics.py:701-716
    import warnings
    import torch
    from deepgpr_amd.preprocessing.model_setup import initialization
    er = torch.full((12, 14), 4.0)
    with warnings.catch_warnings(record=True) as captured:
        warnings.simplefilter("always")
        result = initialization(torch.device("cpu"), er, torch.zeros_like(er), None, torch.zeros((1, 20, 1)),
                                torch.tensor([[[2, 7, 0]]], dtype=torch.int32),
                                torch.tensor([[[6, 7, 0]]], dtype=torch.int32), 0.02, 3.0e-11, 2, 2)
    assert not captured
    assert tuple(result[2:5]) == (16, 18, 1)
