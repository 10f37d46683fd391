import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def gpu():
    """Fail loudly (not skip) when the GPU or libfipi.so is missing."""
    import torch
    import fipi_amd
    assert torch.cuda.is_available(), "GPU test selected but no GPU is visible"
    fipi_amd.lib()
    return torch.device("cuda:0")
