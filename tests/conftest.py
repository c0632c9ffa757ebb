import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.isfile(path):
        pytest.skip(f"golden fixture {name} missing")
    return dict(np.load(path, allow_pickle=False))


def have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.skip("no GPU")
    import i3elvis_amd
    i3elvis_amd.lib()  # raises if the native library is missing
    return True
