import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device and the built HIP library")


@pytest.fixture(scope="session")
def oracle_built():
    """Build the CPU oracle once per session (the checker, never the product)."""
    from oracle import harness
    if not os.path.isfile(harness.ORACLE_LIB):
        harness.build_oracle(reference=True)
    return harness


@pytest.fixture(scope="session")
def cuda_device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")
    from deepgpr_amd.native.loader import get_library
    get_library()  # with a GPU visible, a missing built library must fail, not skip
    return torch.device("cuda:0")
