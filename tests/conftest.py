import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libdroplet_hip.so kernels)")
    config.addinivalue_line("markers", "slow: large inputs")


@pytest.fixture(scope="session")
def device():
    from droplet_amd.renderer import create_device
    dev = create_device()
    yield dev
    dev.destroy()


@pytest.fixture(scope="session")
def renderer(device):
    from droplet_amd.renderer import Renderer
    r = Renderer(device)
    yield r
    r.destroy()
