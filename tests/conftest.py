import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def hip():
    """The HIP library, loaded; fails (does not skip) when it is missing."""
    from pyugks_amd import lib
    L = lib.load()
    if lib.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return lib
