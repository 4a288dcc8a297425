import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and libctl.so")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctl():
    import cutauleaping_amd
    cutauleaping_amd.lib()
    return cutauleaping_amd


@pytest.fixture(scope="session")
def gpu(ctl):
    if ctl.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return ctl
