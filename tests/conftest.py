import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); runs the HIP kernels through the C ABI")


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
