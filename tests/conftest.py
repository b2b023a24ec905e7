import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "webgpu-lvbh-demo_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")


@pytest.fixture(scope="session")
def lb():
    import lbvh
    return lbvh


@pytest.fixture(scope="session")
def ora():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def ctx(lb):
    c = lb.Context(0)
    yield c
    c.close()
