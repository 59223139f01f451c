import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libuhs.so on cuda:0)")


@pytest.fixture(scope="session")
def orc():
    from tests.oracle_py import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def uhs():
    import uxxxxxxxxh_amd as u
    u.lib()
    return u


@pytest.fixture(scope="session")
def gpu(uhs):
    n = uhs.lib().uhs_device_count()
    if n < 1:
        pytest.fail("no GPU visible to libuhs.so (uhs_device_count=%d): %s"
                    % (n, uhs.lib().uhs_last_error().decode()))
    return uhs
