import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "ultimate-kalman_amd", "python"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libukhip's HIP path)")
    config.addinivalue_line("markers", "slow: long CPU case")


@pytest.fixture(scope="session")
def oracle():
    import uk_libs
    return uk_libs.oracle()
