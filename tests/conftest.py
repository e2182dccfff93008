import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libcosma_amd.so")


@pytest.fixture(scope="session")
def handle():
    """One library handle for the GPU tests (device 0)."""
    import torch
    import cosma_amd as ca
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    h = ca.Handle(0)
    yield h
    torch.cuda.synchronize()
    h.close()
