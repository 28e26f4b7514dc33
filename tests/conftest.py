import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsgm_hip on the device)")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def gpu_context():
    from pose_amd import sgm
    if sgm.device_count() == 0:
        pytest.fail("no GPU visible to libsgm_hip")
    ctx = sgm.Context(0)
    yield ctx
    ctx.close()
