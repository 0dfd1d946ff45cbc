import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs via the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    import carpios_amd as ca
    c = ca.Context(0)
    yield c
    c.close()


@pytest.fixture
def planted():
    """Switch on one of the oracle's planted mistakes (oracle.PLANT) for the
    rest of a test; the faithful oracle is restored afterwards."""
    import oracle as O

    def plant(name):
        O.set_plant(name)
    yield plant
    O.set_plant("none")
