import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcie_hip.so on cuda:0)")


@pytest.fixture(scope="session")
def oracle_module():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def torch_gpu():
    import torch

    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    return torch


@pytest.fixture(scope="session")
def hip(torch_gpu):
    """The HIP backend; fails loudly if libcie_hip.so is not built."""
    from cie_amd import _lib, linalg

    _lib.library()
    return linalg
