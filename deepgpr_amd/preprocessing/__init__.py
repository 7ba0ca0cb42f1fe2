"""Input validation, grid spacing / PML parsing, CFL check, model extension."""
from .grid import normalize_grid_spacing, pmlthick_revert  # noqa: F401
from .model_setup import (  # noqa: F401
    NATIVE_FLOAT_DTYPE, PreparedModel, extended_shape_of, initialization, physical_shape_of,
)
from .stability import check_cfl, max_stable_time_step  # noqa: F401
from .wavelets import gaussian, gaussian_derivative, ricker  # noqa: F401
