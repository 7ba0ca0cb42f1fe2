"""ctypes description of the C ABI declared in include/deepgpr_amd.h.

``DgxSolve`` mirrors ``struct dgx_solve``; the ABI-6 tables list the
parameters of the reference-compatible ``forward``/``backward`` exports in
native order (src/DeepGPR/native/abi.py:89-190, src/DeepGPR/lib/deepgpr.h:89-152),
so a header-parsing test can check both sides.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Tuple

from ..config.constants import CPML_FACE_NAMES

DGX_ABI_VERSION = 1

FLOAT_PTR, INT_PTR, VOID_PTR, FLOAT, INT, LONGLONG, CHAR_PTR = (
    "float*", "int*", "void*", "float", "int", "long long", "char*")

_fp = ctypes.POINTER(ctypes.c_float)


class DgxSolve(ctypes.Structure):
    """``struct dgx_solve`` (include/deepgpr_amd.h)."""

    _fields_ = [
        ("abi_version", ctypes.c_int),
        ("fdtd_order", ctypes.c_int),
        ("eps_r_pad", ctypes.c_void_p),
        ("sigma_pad", ctypes.c_void_p),
        ("mu_r_pad", ctypes.c_void_p),
        ("E_saved", ctypes.c_void_p),
        ("R_saved", ctypes.c_void_p),
        ("E", ctypes.c_void_p * 3),
        ("H", ctypes.c_void_p * 3),
        ("update_coeff", ctypes.c_void_p * 6),
        ("phi", ctypes.c_void_p * 24),
        ("pml_coeff", ctypes.c_void_p * 12),
        ("pml", ctypes.c_int * 6),
        ("dt", ctypes.c_float),
        ("nt", ctypes.c_int),
        ("nshot", ctypes.c_int),
        ("nreceiver", ctypes.c_int),
        ("dx", ctypes.c_float),
        ("dy", ctypes.c_float),
        ("dz", ctypes.c_float),
        ("nx_fields", ctypes.c_int),
        ("ny_fields", ctypes.c_int),
        ("nz_fields", ctypes.c_int),
        ("receiver_location", ctypes.c_void_p),
        ("receiver_data", ctypes.c_void_p),
        ("data_grad", ctypes.c_void_p),
        ("receiver_component", ctypes.c_int),
        ("nsource", ctypes.c_int),
        ("source_location", ctypes.c_void_p),
        ("source_waveform", ctypes.c_void_p),
        ("source_component", ctypes.c_int),
        ("grad_source", ctypes.c_void_p),
        ("grad_eps_r", ctypes.c_void_p),
        ("grad_sigma", ctypes.c_void_p),
        ("source_requires_grad", ctypes.c_int),
        ("eps_r_requires_grad", ctypes.c_int),
        ("sigma_requires_grad", ctypes.c_int),
        ("sampling_interval", ctypes.c_int),
        ("fwi_mode", ctypes.c_int),
        ("storage_type", ctypes.c_int),
        ("save_model_history", ctypes.c_int),
        ("use_async_offload", ctypes.c_int),
        ("accumulate_fp64", ctypes.c_int),
    ]


def _phi_names(prefix: str = "") -> Tuple[str, ...]:
    return tuple(f"{prefix}{face}_{field}_phi{k}" for face in CPML_FACE_NAMES
                 for field in ("e", "h") for k in (1, 2))


PHI_PARAMETER_NAMES = _phi_names()
LAMBDA_PHI_PARAMETER_NAMES = _phi_names("lambda_")
_PML_INTS = tuple((f"pml_{f}", INT) for f in CPML_FACE_NAMES)
_PML_COEFFS = tuple((f"{f}_e_coeff", FLOAT_PTR) for f in CPML_FACE_NAMES) + tuple(
    (f"{f}_h_coeff", FLOAT_PTR) for f in CPML_FACE_NAMES)
_MATERIALS = (("eps_r_pad", FLOAT_PTR), ("sigma_pad", FLOAT_PTR), ("mu_r_pad", FLOAT_PTR))
_UPDATE = tuple((n, FLOAT_PTR) for n in ("ce_hist", "ce_curl", "ce_rhs", "ch_hist", "ch_curl", "ch_rhs"))
_TIME = (("dt", FLOAT), ("nt", INT), ("nshot", INT), ("nreceiver", INT),
         ("dx", FLOAT), ("dy", FLOAT), ("dz", FLOAT))
_SIZES = (("nx_fields", INT), ("ny_fields", INT), ("nz_fields", INT))

FORWARD_SIGNATURE = (
    _MATERIALS + (("E_saved", VOID_PTR), ("R_saved", VOID_PTR))
    + tuple((n, FLOAT_PTR) for n in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    + _UPDATE + tuple((n, FLOAT_PTR) for n in PHI_PARAMETER_NAMES) + _PML_INTS + _PML_COEFFS + _TIME
    + (("receiver_location", INT_PTR), ("receiver_data", FLOAT_PTR), ("receiver_component", INT))
    + _SIZES
    + (("nsource", INT), ("source_location", INT_PTR), ("source_waveform", FLOAT_PTR),
       ("source_component", INT), ("sampling_interval", INT), ("fwi_mode", INT),
       ("storage_type", INT), ("save_model_history", INT), ("use_async_offload", INT))
)

BACKWARD_SIGNATURE = (
    _MATERIALS + (("E_saved", VOID_PTR), ("R_saved", VOID_PTR))
    + tuple((n, FLOAT_PTR) for n in ("lambda_ex", "lambda_ey", "lambda_ez",
                                     "lambda_hx", "lambda_hy", "lambda_hz"))
    + _UPDATE + tuple((n, FLOAT_PTR) for n in LAMBDA_PHI_PARAMETER_NAMES) + _PML_INTS + _PML_COEFFS
    + _TIME + _SIZES
    + (("ndata_source", INT), ("receiver_location", INT_PTR), ("data_grad", FLOAT_PTR),
       ("receiver_component", INT), ("nsource", INT), ("source_location", INT_PTR),
       ("source_component", INT), ("grad_source", FLOAT_PTR), ("source_requires_grad", INT),
       ("grad_eps_r", FLOAT_PTR), ("grad_sigma", FLOAT_PTR), ("eps_r_requires_grad", INT),
       ("sigma_requires_grad", INT), ("sampling_interval", INT), ("fwi_mode", INT),
       ("storage_type", INT), ("use_async_offload", INT))
)

CTYPES_BY_KIND: Dict[str, object] = {
    FLOAT_PTR: _fp, INT_PTR: ctypes.POINTER(ctypes.c_int), VOID_PTR: ctypes.c_void_p,
    FLOAT: ctypes.c_float, INT: ctypes.c_int, LONGLONG: ctypes.c_longlong, CHAR_PTR: ctypes.c_char_p,
}

# Every symbol declared in include/deepgpr_amd.h: (parameters, return kind).
NATIVE_FUNCTIONS: Dict[str, Tuple[tuple, object]] = {
    "dgx_abi_version": ((), INT),
    "dgx_last_error": ((), CHAR_PTR),
    "dgx_forward": ((("solve", VOID_PTR), ("stream", VOID_PTR)), INT),
    "dgx_backward": ((("solve", VOID_PTR), ("stream", VOID_PTR)), INT),
    "dgx_forward_noalloc": ((("solve", VOID_PTR), ("stream", VOID_PTR)), INT),
    "dgx_forward_noalloc_scratch_bytes": ((), LONGLONG),
    "dgx_tm2d_tile_rows_tuned": ((("nx_fields", INT), ("ny_fields", INT), ("nshot", INT), ("order", INT),
                                  ("cpml", INT)), INT),
    "dgx_codec_roundtrip": ((("input", FLOAT_PTR), ("encoded", VOID_PTR), ("decoded", FLOAT_PTR),
                             ("count", LONGLONG), ("storage_kind", INT),
                             ("conversion_backend", INT), ("stream", VOID_PTR)), INT),
    "dgx_scratch_bytes": ((("solve", VOID_PTR), ("backward", INT)), LONGLONG),
    "dgx_set_kernel_timing": ((("enable", INT),), None),
    "dgx_kernel_timing": ((("kernel", INT), ("total_ms", VOID_PTR), ("launches", VOID_PTR)), INT),
    "dgx_kernel_name": ((("kernel", INT),), CHAR_PTR),
    "dgx_batch_create": ((("stream", VOID_PTR),), VOID_PTR),
    "dgx_batch_add": ((("batch", VOID_PTR), ("forward", VOID_PTR), ("gradient", VOID_PTR),
                       ("setup_stream", VOID_PTR)), INT),
    "dgx_batch_step": ((("batch", VOID_PTR), ("steps", INT)), INT),
    "dgx_batch_poll": ((("batch", VOID_PTR), ("id", INT)), INT),
    "dgx_batch_active": ((("batch", VOID_PTR),), INT),
    "dgx_batch_destroy": ((("batch", VOID_PTR),), None),
    "deepgpr_supports_error_reporting": ((), INT),
    "deepgpr_last_error": ((), CHAR_PTR),
    "deepgpr_clear_last_error": ((), None),
    "deepgpr_clear_cuda_resource_cache": ((), None),
    "deepgpr_deterministic_adjoint": ((), INT),
    "deepgpr_abi_version": ((), INT),
    "deepgpr_supports_external_pml": ((), INT),
    "deepgpr_supports_pml_material_gradients": ((), INT),
    "deepgpr_supports_int8_wavefield": ((), INT),
    "deepgpr_supports_conversion_backends": ((), INT),
    "deepgpr_supports_int8_reduction_backends": ((), INT),
    "deepgpr_supports_rhs_reconstruction": ((), INT),
    "deepgpr_supports_physical_history": ((), INT),
    "deepgpr_supports_tm2d_fast_path": ((), INT),
    "deepgpr_test_wavefield_conversion": ((("input", FLOAT_PTR), ("encoded", VOID_PTR),
                                           ("decoded", FLOAT_PTR), ("count", LONGLONG),
                                           ("storage_kind", INT), ("conversion_backend", INT)), None),
    "set_fdtd_order": ((("order", INT),), None),
    "forward": (FORWARD_SIGNATURE, None),
    "backward": (BACKWARD_SIGNATURE, None),
}

REQUIRED_SYMBOLS = ("dgx_forward", "dgx_backward", "dgx_abi_version", "forward", "backward",
                    "deepgpr_abi_version")
