import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "the-tessellator_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libtess.so")


@pytest.fixture(scope="session")
def ctx():
    import tess

    c = tess.Context(0)
    yield c
    c.close()


def pytest_runtest_setup(item):
    if item.get_closest_marker("gpu"):
        print(f"\nSTART {item.nodeid}", flush=True)
