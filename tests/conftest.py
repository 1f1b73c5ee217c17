"""pytest configuration: the ``gpu`` marker and shared fixtures.

Tests marked ``gpu`` need a visible GPU and the built ``libck_hip.so``;
everything else runs on the CPU (oracle checks, host logic, the C ABI's
exports and argument validation).
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def gpu():
    """The GPU device; fails (not skips) when a gpu-marked test has no GPU."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test without a visible GPU")
    from commkit_amd import _native

    _native.lib()  # raises when libck_hip.so is missing
    return torch.device("cuda", 0)


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
