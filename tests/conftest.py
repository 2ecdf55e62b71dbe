import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long-running full-size check")


def _have_gpu():
    try:
        import scip_amd
        return scip_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    import scip_amd
    if scip_amd.device_count() == 0:
        pytest.fail("no HIP device visible; gpu tests need an MI355X")
    ctx = scip_amd.Context(0)
    yield ctx
    ctx.close()
