import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O.oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import oracle as O
    r = O.reference()
    if r is None:
        pytest.skip("oracle/_ref (the compiled reference) is not built here")
    return r
