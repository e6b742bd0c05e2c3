import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    _oracle.load()
    return _oracle


@pytest.fixture(scope="session")
def engine_factory():
    """Creates engines on device 0; fails loudly if the HIP library is missing."""
    from gpusharesat_amd.engine import Engine
    made = []

    def make():
        e = Engine(0)
        made.append(e)
        return e

    yield make
    for e in made:
        e.close()
