"""Shared test configuration.

Marker ``gpu``: the test needs a real MI355X and the built HIP library; the
CPU suite runs with ``-m "not gpu"``.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libxpbd_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc

    if not orc.LIB.exists():
        orc.build()
    return orc


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a visible GPU")
    return torch.device("cuda:0")
