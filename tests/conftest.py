import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests must run on a GPU and load the HIP library."""
    import torch  # device memory for device-pointer tests
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected without a visible GPU")
    from sgpp_amd import capi
    return capi
