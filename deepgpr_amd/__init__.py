"""deepgpr_amd: the DeepGPR FDTD hot path, MI355X-native.

Drop-in for the reference ``DeepGPR.compute`` / ``checkpoint_initial_field``
API (src/DeepGPR/__init__.py:32-79). The forward solver, the exact discrete
adjoint, CPML, sources/receivers and the wavefield-history codec run as
hand-written HIP kernels for gfx950 in ``libdeepgpr_amd.so`` (C ABI in
``include/deepgpr_amd.h``); PyTorch provides tensors and autograd only.
"""
from __future__ import annotations

from . import config, utils, wavelet  # noqa: F401
from .config.constants import SPEED_OF_LIGHT as c  # noqa: F401
from .native.loader import (  # noqa: F401
    clear_cuda_resource_cache, get_deepgpr_lib, get_deepgpr_library_path, get_library, library_path,
    set_library_fdtd_order,
)
from .preprocessing.grid import pmlthick_revert  # noqa: F401
from .preprocessing.model_setup import initialization  # noqa: F401
from .preprocessing.stability import check_cfl, max_stable_time_step  # noqa: F401
from .preprocessing.wavelets import gaussian, gaussian_derivative, ricker  # noqa: F401
from .solver.autograd import DeepGPR, check_tensors_for_nan_inf  # noqa: F401
from .solver.fields import checkpoint_initial_field, create_or_separate, zero_field  # noqa: F401
from .solver.memory import estimate_compute_memory  # noqa: F401
from .solver.modeling import compute  # noqa: F401
from .solver.pml import (  # noqa: F401
    CFS, CFSParameter, build_pml_coeffs, build_pml_phi, buildpmlcoeffs, calculate_pml_update_coeffs, e0, m0,
)
from .solver.sampling import recommended_sampling_interval, source_max_frequency  # noqa: F401
from .solver.segments import plan_segments, segmented_compute  # noqa: F401
from .solver.storage import decompress_wavefield_history  # noqa: F401
from .utils.logger import configure_logging, get_logger  # noqa: F401
from .utils.exceptions import (  # noqa: F401
    CFLConditionError, DeepGPRError, NativeLibraryError, NativeLibraryNotFoundError, NonFiniteTensorError,
)

__version__ = "0.1.0"
