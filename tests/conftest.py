import os
import sys

import numpy as np
import pytest

# torch ships its own HIP runtime under the same soname as ROCm's; loaded
# first, it is the one the library binds to as well, and torch's device
# enumeration works in tests that hand torch-allocated buffers to the library.
try:
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is part of the image
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session", params=["mfma", "valu"])
def ctx(request):
    """A GPU context per lead-contraction path: every GPU test runs on both
    the matrix-core and the vector-ALU kernel."""
    from ekgsim_amd import Context
    c = Context(0)
    c.set_kernel(request.param)
    yield c
    c.close()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
