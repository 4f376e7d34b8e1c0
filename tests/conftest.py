import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libgenfire kernels)")
    config.addinivalue_line("markers", "slow: longer CPU oracle checks")


@pytest.fixture(scope="session")
def ctx():
    from genfire_amd import _lib
    return _lib.Context(0)


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    return oracle
