"""Loading and calling the HIP library (libdeepgpr_amd.so)."""
from .loader import get_library, library_path, require_library  # noqa: F401
