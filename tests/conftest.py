import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tristan-ff_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through libtff.so)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    """The CPU oracle is test infrastructure: compile it if it is missing."""
    from oracle import oracle as orc
    if not os.path.exists(os.path.join(orc.HERE, "_build", "liboracle_f64.so")):
        orc.build(ref=False)
    yield
