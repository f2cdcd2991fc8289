import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session", autouse=True)
def _torch_device_first(request):
    """torch's bundled HIP runtime initialises before libtrjhip's (the two are
    separate runtimes over one HSA runtime; torch finds no device when the
    library's runtime came up first in the process)."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        if torch.cuda.device_count() > 0:
            torch.cuda.init()
    yield
