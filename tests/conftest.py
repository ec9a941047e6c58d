import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")
    # the CPU oracle is test infrastructure; build it on first use if it is absent
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    """The GPU device context; GPU tests fail loudly if the HIP library is missing."""
    if not _gpu_available():
        pytest.skip("no ROCm GPU in this environment")
    from miccai2016_amd.device import default_context
    return default_context(0)


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
