import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def O():
    import __graft_entry__ as ge
    return ge.load_oracle()


@pytest.fixture(scope="session")
def C():
    import __graft_entry__ as ge
    return ge.load_csaw()
