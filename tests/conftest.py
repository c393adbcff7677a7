import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a MI355X (gfx950) GPU and the built libm25x.so")


@pytest.fixture(scope='session')
def gpu():
    """The GPU tests' guard: a device must be present and the library built (no fallback exists)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mom25d_amd import device as dv
    dv.device()
    return dv
