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
    """The HIP library, rebuilt first when it is missing or was built from
    sources other than the ones in the tree (compared by content hash)."""
    import time

    from vortegrid_amd import build, vg as V

    if not build.is_current():
        print(f"[vg] library missing or stale (built {build.built_hash()}, sources "
              f"{build.source_hash()}): rebuilding", flush=True)
        build.build()
    print(f"[vg] {V.LIB_PATH} built {time.ctime(os.path.getmtime(V.LIB_PATH))}, "
          f"source hash {build.built_hash()}", flush=True)
    V.require_gpu()
    return V
