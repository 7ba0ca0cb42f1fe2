"""Loading and calling the HIP library (libdeepgpr_amd.so)."""
from .loader import (  # noqa: F401
    clear_cuda_resource_cache, get_deepgpr_lib, get_deepgpr_library_path, get_library, library_path,
    library_supports, require_library, set_library_fdtd_order,
)
