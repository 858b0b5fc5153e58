import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """The GPU test context: torch with a visible device and the native library loaded."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import naja_amd

    naja_amd.load_library()  # raises if the HIP library was not built
    return torch
