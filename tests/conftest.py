import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def eng():
    import fftw3_amd
    fftw3_amd.lib()
    fftw3_amd.init()
    return fftw3_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
