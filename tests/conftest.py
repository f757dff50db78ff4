import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")


@pytest.fixture(scope="session")
def gpu():
    """Fail loudly (never skip) when a gpu-marked test runs without a GPU or without the library."""
    import torch
    assert torch.cuda.is_available(), "gpu test run without a visible GPU"
    from pylians_amd import _lib
    _lib.lib()
    return torch.device("cuda")
