import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and the built libhplx.so")


@pytest.fixture(scope="session")
def ctx():
    """One library context for all GPU tests of the session."""
    import rochpl_amd as rp
    print("\n[gpu] opening hplx context on device 0", flush=True)
    c = rp.Context(0)
    yield c
    c.close()
