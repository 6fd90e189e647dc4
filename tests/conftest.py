import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def rf():
    import redfox_amd

    if redfox_amd.device_count() < 1:
        raise RuntimeError("GPU test run without a visible HIP device")
    return redfox_amd


@pytest.fixture(scope="session")
def ctx(rf):
    c = rf.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
