import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libcompas_amd.so")


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests run the HIP library and nothing else: no GPU or no library is a failure."""
    import torch
    assert torch.cuda.is_available(), "GPU test started without a visible GPU"
    from compas_amd import capi
    capi.load()   # raises when libcompas_amd.so is missing
    return torch.device("cuda:0")
