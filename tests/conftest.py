import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import Reference, reference_available
    if not reference_available():
        pytest.skip("compiled reference (oracle/_ref) not built here")
    return Reference()


@pytest.fixture(scope="session")
def gpu():
    """The product library with a GPU behind it; fails (not skips) without one."""
    import s2tc_amd
    L = s2tc_amd.lib()
    n = L.s2tc_amd_device_count()
    assert n > 0, "no GPU visible to the HIP runtime"
    return s2tc_amd
