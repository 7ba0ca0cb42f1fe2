"""Solver host layer: CPML, fields, history storage, autograd and compute."""
from .autograd import DeepGPR, SolverConfig, SolverTensors  # noqa: F401
from .fields import checkpoint_initial_field, create_or_separate, zero_field  # noqa: F401
from .memory import estimate_compute_memory  # noqa: F401
from .modeling import compute  # noqa: F401
from .pml import (  # noqa: F401
    CFS, CFSParameter, build_pml_coeffs, build_pml_phi, buildpmlcoeffs, calculate_pml_update_coeffs,
    pml_face_descriptors, pml_phi_element_count,
)
from .sampling import recommended_sampling_interval, source_max_frequency  # noqa: F401
from .segments import plan_segments, segmented_compute  # noqa: F401
from .storage import WavefieldStorageConfig, decompress_wavefield_history, int8_history_layout  # noqa: F401
