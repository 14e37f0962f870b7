import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as ge
    return ge.import_package()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
