import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device 0)')


@pytest.fixture(scope='session')
def oracle():
    from oracle import oracle as O
    return O


@pytest.fixture(scope='session')
def hip():
    """The HIP library, initialised on device 0 (GPU tests only)."""
    from lentils_amd import backend as B
    B.check(B.lib().lt_init(0))
    return B
