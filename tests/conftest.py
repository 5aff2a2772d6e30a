import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the library on cuda:0; the HIP path must be the one that runs."""
    import torch
    import graphbench_amd as gb
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected on a machine without a HIP device")
    gb.init(0)
    torch.zeros(1, device="cuda")
    return gb
