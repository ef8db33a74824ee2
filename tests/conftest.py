import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "rmsd_reference_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden():
    return dict(np.load(GOLDEN))


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: a gpu-marked test on a machine without a GPU is a
    failure, not a skip (the HIP path must be the one that runs)."""
    import torch

    assert torch.cuda.is_available(), "gpu test needs a visible GPU"
    from mdtraj_amd import _lib

    _lib.load()
    return torch
