import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    O.set_fault(O.FAULT_NONE)
    return O


@pytest.fixture(scope="session")
def spg():
    import spgemm_amd
    spgemm_amd.lib()
    return spgemm_amd


@pytest.fixture(scope="session")
def gpu_ctx(spg):
    """One context for all GPU tests (one process, one device)."""
    if spg.device_count() < 1:
        pytest.fail("GPU test collected but no device is visible")
    ctx = spg.Context(0)
    ctx.defaults = ctx.get_config()
    yield ctx
    ctx.close()


@pytest.fixture(autouse=True)
def _fresh_config(request):
    """Every test that uses the shared context starts from a new context's
    configuration, so no setting leaks from one test into the next."""
    if "gpu_ctx" in request.fixturenames:
        ctx = request.getfixturevalue("gpu_ctx")
        ctx.set_config(**ctx.defaults)
    yield
