import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def vg():
    """The HIP library; built here only when no built copy exists."""
    from vortegrid_amd import build, vg as V

    if not os.path.exists(V.LIB_PATH):
        build.build()
    V.require_gpu()
    return V
