import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One device context shared by the GPU tests (the HIP library must load;
    there is no fallback)."""
    from rchem_amd.basis import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
