import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    return pyoracle.load()


@pytest.fixture(scope="session")
def gpu():
    """The product binding; a GPU test fails (does not skip) if the HIP library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected on a machine without a GPU")
    import dtrt_amd.dtrt as G
    return G
