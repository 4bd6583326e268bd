import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "nbody-sim_amd"):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libnbx.so")


@pytest.fixture(scope="session")
def gpu():
    """The nbx module with its HIP library loaded; fails loudly if unusable."""
    import torch
    import nbx
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    nbx.lib()
    return nbx
