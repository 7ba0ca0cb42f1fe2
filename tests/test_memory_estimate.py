"""estimate_compute_memory under the reference's keywords and result keys
(src/DeepGPR/solver/memory.py:30-246; its test tests/test_numerics.py:85-115),
priced for this build (no GPU needed)."""
import inspect
import math

import pytest
import torch

import deepgpr_amd as dg
from deepgpr_amd.solver.storage import int8_history_layout
from deepgpr_amd.utils.exceptions import NativeLibraryError

# keyword-only parameters of the reference function, in order (memory.py:30-53)
REFERENCE_KEYWORDS = ["device", "nx", "ny", "nz", "nt", "nstep", "nsr", "nrx", "source_waveforms", "pml", "mode",
                      "sampling_interval", "storage_dtype", "use_async_offload", "er_requires_grad",
                      "se_requires_grad", "save_wavefield_history", "wavefield_compression",
                      "compression_block_size", "reconstruct_rhs", "history_region"]
# result keys of the reference function (memory.py:220-246)
REFERENCE_KEYS = {"model_and_padded_materials", "electric_and_magnetic_fields", "cpml_auxiliary_fields",
                  "cpml_coefficients", "source_and_locations", "saved_gradient_wavefields", "final_e_frame",
                  "packed_history_per_quantity", "fdtd_update_coefficients", "receiver_data", "material_gradients",
                  "adjoint_fields_and_cpml", "receiver_adjoint", "adjoint_stencil_weights",
                  "low_precision_exact_snapshot", "cuda_transfer_buffers", "estimated_device_peak",
                  "estimated_host_peak", "recommended_device_capacity", "recommended_host_capacity", "nt_saved",
                  "components_saved", "history_shape", "reconstruct_rhs", "effective_async_offload"}

CUDA = torch.device("cuda")


def _common(**over):
    args = dict(device=CUDA, nx=13, ny=18, nz=1, nt=5, nstep=2, nsr=1, nrx=3, source_waveforms=1, pml=[0] * 6,
                mode=2, sampling_interval=1, storage_dtype=torch.float32, use_async_offload=False,
                er_requires_grad=True, se_requires_grad=True)
    args.update(over)
    return args


def test_signature_and_keys_are_the_reference_ones():
    params = inspect.signature(dg.estimate_compute_memory).parameters
    names = [n for n, p in params.items() if p.kind is inspect.Parameter.KEYWORD_ONLY]
    assert names[:len(REFERENCE_KEYWORDS)] == REFERENCE_KEYWORDS
    assert names[len(REFERENCE_KEYWORDS):] == ["gradient_accumulator"]
    assert not [n for n, p in params.items() if p.kind is not inspect.Parameter.KEYWORD_ONLY]
    assert set(dg.estimate_compute_memory(**_common())) == REFERENCE_KEYS


def test_int8_estimate_includes_fp32_scales():
    # the reference's own case (tests/test_numerics.py:85-115): E and R INT8
    # histories, each packed with its FP32 block scales
    fp32 = dg.estimate_compute_memory(**_common())
    int8 = dg.estimate_compute_memory(**_common(), wavefield_compression="int8", compression_block_size=(8, 8, 1))
    one = int8_history_layout((5, 2, 13, 18, 1), (8, 8, 1))["packed_bytes"]
    assert int8["saved_gradient_wavefields"] == 2 * one
    assert int8["packed_history_per_quantity"] == one
    assert int8["saved_gradient_wavefields"] < fp32["saved_gradient_wavefields"]


def test_history_rules():
    frame = 2 * 13 * 18 * 1 * 4  # one FP32 frame of both shots, one component
    auto = dg.estimate_compute_memory(**_common())  # FP32 at S 1: E-only plus the final frame
    assert auto["reconstruct_rhs"] and auto["final_e_frame"] == frame
    assert auto["saved_gradient_wavefields"] == 5 * frame + frame
    stored = dg.estimate_compute_memory(**_common(reconstruct_rhs=False))
    assert stored["saved_gradient_wavefields"] == 2 * 5 * frame and stored["final_e_frame"] == 0
    bf16 = dg.estimate_compute_memory(**_common(storage_dtype=torch.bfloat16, sampling_interval=2))
    assert not bf16["reconstruct_rhs"] and bf16["nt_saved"] == 3
    assert bf16["saved_gradient_wavefields"] == 2 * 3 * frame // 2
    forward = dg.estimate_compute_memory(**_common(er_requires_grad=False, se_requires_grad=False))
    assert forward["saved_gradient_wavefields"] == 5 * frame  # E only, no final frame
    assert forward["adjoint_fields_and_cpml"] == forward["receiver_adjoint"] == forward["material_gradients"] == 0
    none = dg.estimate_compute_memory(**_common(save_wavefield_history=False))
    assert none["saved_gradient_wavefields"] == 0 and none["adjoint_fields_and_cpml"] == 0


def test_physical_region_and_components():
    pml = [4, 5, 3, 4, 2, 3]
    full = dg.estimate_compute_memory(**_common(nx=30, ny=32, nz=28, pml=pml, mode=3, reconstruct_rhs=False))
    phys = dg.estimate_compute_memory(**_common(nx=30, ny=32, nz=28, pml=pml, mode=3, reconstruct_rhs=False,
                                                history_region="physical"))
    assert full["history_shape"] == (30, 32, 28) and phys["history_shape"] == (21, 25, 23)
    assert full["components_saved"] == 3
    assert phys["saved_gradient_wavefields"] == 2 * 3 * 5 * 2 * 21 * 25 * 23 * 4
    tiles = dg.estimate_compute_memory(**_common(nx=30, ny=32, nz=28, pml=pml, mode=3, wavefield_compression="int8",
                                                 history_region="physical"))
    # the physical box [4, 25) x [3, 28) x [2, 25) widened to whole 4x4x4 tiles
    assert tiles["history_shape"] == (24, 28, 28)


def test_this_builds_device_terms():
    nx, ny, nz, nstep = 20, 22, 24, 3
    fields = (nx + 1) * (ny + 1) * (nz + 1)
    args = _common(nx=nx, ny=ny, nz=nz, nstep=nstep, mode=3)
    f64 = dg.estimate_compute_memory(**args)
    f32 = dg.estimate_compute_memory(**args, gradient_accumulator="float32")
    # twelve FP32 channel weights per cell and shot for the 3D adjoint
    assert f64["adjoint_stencil_weights"] == 48 * nstep * fields
    assert f64["material_gradients"] - f32["material_gradients"] == 16 * nx * ny * nz
    assert f64["cuda_transfer_buffers"] == f64["low_precision_exact_snapshot"] == 0
    offload = dg.estimate_compute_memory(**dict(args, use_async_offload=True))
    assert not offload["effective_async_offload"] and offload == f64  # histories stay in HBM
    assert f64["estimated_host_peak"] == f64["recommended_host_capacity"] == 0
    assert f64["recommended_device_capacity"] == int(f64["estimated_device_peak"] * 1.2)
    parts = sum(f64[k] for k in ("model_and_padded_materials", "electric_and_magnetic_fields",
                                 "cpml_auxiliary_fields", "cpml_coefficients", "source_and_locations",
                                 "saved_gradient_wavefields", "fdtd_update_coefficients", "receiver_data",
                                 "material_gradients", "adjoint_fields_and_cpml", "receiver_adjoint",
                                 "adjoint_stencil_weights"))
    assert f64["estimated_device_peak"] == parts  # the backward holds every term at once
    tm = dg.estimate_compute_memory(**_common())
    assert tm["adjoint_stencil_weights"] == 0  # the 2D kernels recompute their weights


def test_cpu_device_has_no_backend():
    with pytest.raises(NativeLibraryError):
        dg.estimate_compute_memory(**_common(device=torch.device("cpu")))
    assert math.isclose(dg.estimate_compute_memory(**_common(device="cuda:0"))["estimated_device_peak"],
                        dg.estimate_compute_memory(**_common())["estimated_device_peak"])


def test_estimate_scales_with_physical_cells():
    # restates tests/test_history_region.py:140-153 (on a HIP device: this
    # build has no CPU backend)
    common = dict(device=CUDA, nx=30, ny=26, nz=1, nt=50, nstep=2, nsr=1, nrx=3, source_waveforms=1,
                  pml=[3, 5, 4, 2, 0, 0], mode=2, sampling_interval=1, storage_dtype=torch.float32,
                  use_async_offload=False, er_requires_grad=False, se_requires_grad=False)
    extended = dg.estimate_compute_memory(**common)
    physical = dg.estimate_compute_memory(history_region="physical", **common)
    assert physical["history_shape"] == (22, 20, 1)
    assert physical["saved_gradient_wavefields"] * 30 * 26 == extended["saved_gradient_wavefields"] * 22 * 20


def test_one_history_and_the_final_frame():
    # restates tests/test_rhs_history.py:169-186; where the reference keeps an
    # exact FP32 frame beside a low-precision history, this build forms R^n
    # from the exact E^n in the E update's registers and keeps none
    common = dict(device=CUDA, nx=20, ny=24, nz=1, nt=50, nstep=2, nsr=1, nrx=3, source_waveforms=1,
                  pml=[3, 3, 3, 3, 0, 0], mode=2, sampling_interval=1, use_async_offload=False,
                  er_requires_grad=True, se_requires_grad=False)
    one_history = 50 * 2 * 20 * 24 * 4
    e_only = dg.estimate_compute_memory(storage_dtype=torch.float32, **common)
    pair = dg.estimate_compute_memory(storage_dtype=torch.float32, reconstruct_rhs=False, **common)
    half = dg.estimate_compute_memory(storage_dtype=torch.float16, **common)
    assert e_only["reconstruct_rhs"]
    assert e_only["final_e_frame"] == 2 * 20 * 24 * 4
    assert e_only["saved_gradient_wavefields"] == one_history + 2 * 20 * 24 * 4
    assert pair["saved_gradient_wavefields"] == 2 * one_history
    assert not half["reconstruct_rhs"] and half["saved_gradient_wavefields"] == one_history
    assert half["low_precision_exact_snapshot"] == 0


def test_cpml_term_matches_the_allocated_tensors():
    # restates tests/test_numerics.py:345-369
    from deepgpr_amd.solver.pml import build_pml_phi, buildpmlcoeffs
    nx, ny, nz, nstep = 10, 12, 14, 3
    pml = [2, 3, 4, 2, 3, 1]
    er = torch.full((nx, ny, nz), 4.0)
    mr = torch.ones((nx + 1, ny + 1, nz + 1))
    faces = buildpmlcoeffs(er, mr, 2.0e-11, 0.02, nx, ny, nz, torch.tensor(pml, dtype=torch.int32),
                           torch.device("cpu"), torch.float32)[:6]
    tensors = build_pml_phi(*faces, nstep, None, torch.device("cpu"))
    est = dg.estimate_compute_memory(**_common(nx=nx, ny=ny, nz=nz, nstep=nstep, pml=pml, mode=3))
    assert est["cpml_auxiliary_fields"] == 4 * sum(t.numel() for t in tensors)


def _config(**over):
    from deepgpr_amd.solver.autograd import SolverConfig
    from deepgpr_amd.solver.storage import WavefieldStorageConfig
    args = dict(device=CUDA, dtype=torch.float32, grid_spacing=(0.02, 0.02, 0.02), dt=3e-11, nx=18, ny=22, nz=1,
                nt=80, nstep=1, nsr=1, nrx=1, pmlthick=torch.tensor([3, 3, 3, 3, 0, 0], dtype=torch.int32),
                source_direction=2, receiver_component=2, sampling_interval=2,
                storage=WavefieldStorageConfig(torch.float16, "none", None, "current", "current"),
                save_wavefield_history=True, use_async_offload=False, fdtd_order=4, mode=2, debug=False,
                tm2d_fast_path=True)
    args.update(over)
    return SolverConfig(**args)


def test_preview_reports_the_estimate():
    # the labels the reference's preview test looks for (tests/test_numerics.py:262-299),
    # with the HIP device in place of the CPU
    from deepgpr_amd.solver.memory import format_compute_preview
    text = format_compute_preview(_config(), source_waveforms=1, er_requires_grad=True, se_requires_grad=True)
    for line in ("=== DeepGPR compute preview ===", "dx / dy / dz", "FDTD order / gradient mode: 4 / 2",
                 "saved E_saved and R_saved wavefields", "adjoint fields and CPML",
                 "forward wavefield save directory: disabled", "estimated peak HIP device memory",
                 "recommended device capacity with 20% margin", "=== End DeepGPR compute preview ==="):
        assert line in text, line
    est = dg.estimate_compute_memory(device=CUDA, nx=18, ny=22, nz=1, nt=80, nstep=1, nsr=1, nrx=1,
                                     source_waveforms=1, pml=[3, 3, 3, 3, 0, 0], mode=2, sampling_interval=2,
                                     storage_dtype=torch.float16, use_async_offload=False, er_requires_grad=True,
                                     se_requires_grad=True, reconstruct_rhs=False)
    assert f"estimated peak HIP device memory: {est['estimated_device_peak'] / 2**20:.2f} MiB" in text
    saved = format_compute_preview(_config(), source_waveforms=1, er_requires_grad=False, se_requires_grad=False,
                                   save_directory="/tmp/run")
    assert "forward wavefield save directory: /tmp/run" in saved
