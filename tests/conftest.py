import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ora():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def pla_session():
    """One PLA_Init / 1x1 template for the whole GPU test session."""
    from plapack_amd import pla
    with pla.Session(nb_distr=128) as s:
        yield s
