import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pg-strom_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the HIP device once per test session."""
    from pgstrom import xpuhip
    xpuhip.init(0)
    return xpuhip


@pytest.fixture(params=[2, 1], ids=["interp", "jit"])
def jit(request):
    """Both kernels of a plan: the op interpreter (XPUHIP_JIT_OFF) and the
    kernel specialized for the plan with hiprtc (XPUHIP_JIT_ON)."""
    return request.param
