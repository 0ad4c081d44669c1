import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "cuda-to-sycl-nbody_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP device)")


@pytest.fixture(scope="session")
def nb():
    import nbody_amd

    nbody_amd.load()
    return nbody_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu(nb):
    """Skips only when no GPU is visible; on a machine with a GPU the HIP
    library must load and run (no fallback)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    assert nb.device_count() >= 1, "torch sees a GPU but libnbody_amd does not"
    return torch.device("cuda:0")
