import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); parity tests proper")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc
    orc.build()
    return orc


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) without a GPU: GPU tests only run where one is expected."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible ROCm device"
    return torch.device("cuda:0")
