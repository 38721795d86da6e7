import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP binding; GPU tests fail (not skip) if the library is missing."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import torch_amd

    torch_amd.lib()
    return torch_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
