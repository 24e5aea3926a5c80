import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
os.environ.setdefault("FFTGAL_VERBOSE", "0")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch


@pytest.fixture(scope="session")
def lib():
    import fftgal_amd
    return fftgal_amd.load()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle
