import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libjaxpm_hip.so")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is visible; returns the torch device."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxpm_amd import _lib
    return _lib.device()


def rng(seed):
    return np.random.default_rng(seed)


# The reference's test configurations (tests/conftest.py:37-49 of JaxPM).
SIM_CONFIGS = [((32, 32, 32), (256., 256., 256.)), ((32, 32, 64), (256., 256., 512.))]
LPT_SCALE_FACTORS = [0.1, 0.2]
