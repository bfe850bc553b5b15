import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("cuda-riesel-sieve_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
R745 = os.path.join(GOLDEN, "sr_745.abcd")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def rs():
    import pyrsieve

    pyrsieve.lib()
    return pyrsieve


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(rs):
    if rs.device_count() < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return 0
