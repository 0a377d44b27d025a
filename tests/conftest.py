import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def hip_ctx():
    """One library context for the whole GPU session."""
    import torch
    from cfdarcov2_amd import capi
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    ctx = capi.Context(0)
    yield ctx
    torch.cuda.synchronize()
    ctx.close()
