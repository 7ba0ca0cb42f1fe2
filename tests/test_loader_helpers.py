"""The reference's top-level loader and logging helpers
(src/DeepGPR/__init__.py:36-45, native/loader.py:222-293, utils/logger.py:36-64,
wavelet.py)
under the same names, with the same argument checks and errors."""
import io
import logging

import pytest

import deepgpr_amd as dg
from deepgpr_amd.utils.exceptions import NativeLibraryError

NAMES = ("clear_cuda_resource_cache", "get_deepgpr_lib", "get_deepgpr_library_path", "set_library_fdtd_order",
         "configure_logging", "get_logger")


@pytest.mark.parametrize("name", NAMES)
def test_exported_at_top_level(name):
    assert callable(getattr(dg, name))


def test_library_for_device_kinds():
    from deepgpr_amd.native.loader import library_path
    assert dg.get_deepgpr_library_path("cuda") == library_path()
    assert dg.get_deepgpr_lib("cuda:0") is dg.get_library()
    with pytest.raises(NativeLibraryError):  # no CPU backend in this build
        dg.get_deepgpr_lib("cpu")
    with pytest.raises(ValueError):
        dg.get_deepgpr_lib("meta")


def test_set_library_fdtd_order_checks_the_order():
    lib = dg.get_deepgpr_lib("cuda")
    for bad in (0, 3, 6, 16):
        with pytest.raises(ValueError):
            dg.set_library_fdtd_order(lib, bad)
    for order in (4, 8, 2):  # host-side state only; leaves the default order 2 set
        dg.set_library_fdtd_order(lib, order)


def test_clear_cuda_resource_cache_refuses_a_cpu_device():
    with pytest.raises(ValueError):
        dg.clear_cuda_resource_cache("cpu")


def test_configure_logging_replaces_its_handler():
    buf = io.StringIO()
    logger = dg.configure_logging("DEBUG", "%(name)s|%(message)s", stream=buf)
    try:
        dg.configure_logging("DEBUG", "%(name)s|%(message)s", stream=buf)
        ours = [h for h in logger.handlers if isinstance(h, logging.StreamHandler)
                and not isinstance(h, logging.NullHandler)]
        assert len(ours) == 1
        dg.get_logger("solver").debug("one line")
        assert buf.getvalue() == "DeepGPR.solver|one line\n"
        assert not set(ours) & set(logging.getLogger().handlers)  # the root logger is left alone
    finally:
        for h in list(logger.handlers):
            if getattr(h, "_deepgpr_amd_console", False):
                logger.removeHandler(h)
        logger.setLevel(logging.NOTSET)


@pytest.mark.gpu
def test_clear_cuda_resource_cache_on_the_device(cuda_device):
    print("[gpu loader] clear_cuda_resource_cache", flush=True)
    dg.clear_cuda_resource_cache()
    dg.clear_cuda_resource_cache(0)
    dg.clear_cuda_resource_cache(cuda_device)


def test_wavelet_namespace_of_the_quick_start():
    assert dg.wavelet.ricker is dg.ricker
    w = dg.wavelet.ricker(2e8, 2000, 3e-11, 5e-9)
    assert w.shape == (2000,) and float(w.abs().max()) == pytest.approx(1.0, abs=1e-3)


SUBPACKAGE_NAMES = {
    # src/DeepGPR/solver/__init__.py:2-18 (format_compute_preview, whose
    # arguments are compute()'s resolved state, is not exported here)
    "deepgpr_amd.solver": ["DeepGPR", "SolverConfig", "SolverTensors", "WavefieldStorageConfig", "CFS",
                           "CFSParameter", "build_pml_coeffs", "build_pml_phi", "buildpmlcoeffs",
                           "calculate_pml_update_coeffs", "checkpoint_initial_field", "compute", "create_or_separate",
                           "decompress_wavefield_history", "estimate_compute_memory", "int8_history_layout",
                           "pml_face_descriptors", "pml_phi_element_count", "recommended_sampling_interval",
                           "source_max_frequency", "zero_field"],
    # src/DeepGPR/preprocessing/__init__.py:2-11 (without morlet and sine_burst)
    "deepgpr_amd.preprocessing": ["NATIVE_FLOAT_DTYPE", "PreparedModel", "check_cfl", "extended_shape_of",
                                  "gaussian", "gaussian_derivative", "initialization", "max_stable_time_step",
                                  "normalize_grid_spacing", "physical_shape_of", "pmlthick_revert", "ricker"],
}


@pytest.mark.parametrize("module", sorted(SUBPACKAGE_NAMES))
def test_subpackages_export_the_reference_names(module):
    import importlib
    mod = importlib.import_module(module)
    assert [n for n in SUBPACKAGE_NAMES[module] if not hasattr(mod, n)] == []


def test_shape_helpers_are_inverse():
    from deepgpr_amd.preprocessing import extended_shape_of, physical_shape_of
    pml = [4, 5, 3, 4, 2, 3]
    assert extended_shape_of((20, 21, 22), pml) == (29, 28, 27)
    assert physical_shape_of(extended_shape_of((20, 21, 22), pml), pml) == (20, 21, 22)
