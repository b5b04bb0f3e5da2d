import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def gpu():
    """Device fixture for GPU tests: fails (does not skip) without a HIP device, and makes
    sure libhfx itself is what runs."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    from holoflow_amd import _lib

    _lib.lib()
    torch.cuda.set_device(0)
    return torch.device("cuda:0")


def rel_l2(x, ref):
    import numpy as np

    x = np.asarray(x)
    ref = np.asarray(ref)
    dt = np.complex128 if (np.iscomplexobj(x) or np.iscomplexobj(ref)) else np.float64
    x = x.astype(dt)
    ref = ref.astype(dt)
    d = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (d if d > 0 else 1.0))
