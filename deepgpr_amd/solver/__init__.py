"""Solver host layer: CPML, fields, history storage, autograd and compute."""
from .autograd import DeepGPR, SolverConfig, SolverTensors  # noqa: F401
from .fields import checkpoint_initial_field, create_or_separate, zero_field  # noqa: F401
from .modeling import compute  # noqa: F401
from .pml import CFS, CFSParameter, build_pml_coeffs, build_pml_phi, calculate_pml_update_coeffs  # noqa: F401
from .sampling import recommended_sampling_interval, source_max_frequency  # noqa: F401
from .storage import decompress_wavefield_history  # noqa: F401
