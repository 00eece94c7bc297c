import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and master_amd/libdsmc.so")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden():
    return dict(np.load(GOLDEN, allow_pickle=False))


@pytest.fixture(scope="session")
def world32():
    from oracle import binding as B
    return B.packed_spheres(32, 4.0, 0.6, 1)


def gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips only when no GPU is visible; on a GPU box a missing library fails."""
    if not gpu_present():
        pytest.skip("no HIP device visible")
    from master_amd import dsmc
    dsmc.lib()  # raises if the HIP library is missing: no silent fallback
    return dsmc
