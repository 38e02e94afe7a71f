import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); calls libdrawhip through its C ABI")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def renderer():
    """One libdrawhip context for the GPU tests (fails loudly when the library
    or the device is missing: there is no CPU fallback)."""
    from draw_amd.render import Renderer
    r = Renderer(0)
    yield r
    r.close()
