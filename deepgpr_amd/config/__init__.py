"""Constants and defaults of the deepgpr_amd API."""
from . import constants, defaults  # noqa: F401
