import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP path")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def dev():
    """cuda:0 with the HIP library loaded; fails (never skips) when either is missing."""
    import torch
    import dumbray_amd
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    dumbray_amd.lib()
    print(f"[gpu] device={torch.cuda.get_device_name(0)}", flush=True)
    return torch.device("cuda:0")
