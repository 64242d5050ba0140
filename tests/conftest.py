import os
import sys

import numpy as np
import pytest
import torch  # noqa: F401  (load torch's HIP runtime before any native library)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the native HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is visible; returns the package."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import hartree_fock_cpp_amd as hfx

    return hfx
