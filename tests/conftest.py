import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def orc():
    """The C oracle (built on first use)."""
    from oracle import oracle as O
    O.lib()
    return O


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ws():
    """The HIP library, loaded through ctypes; GPU tests fail loudly without it."""
    import windstencil_amd as W
    W.lib()          # raises if the shared library is missing
    return W
