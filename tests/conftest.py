import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context on cuda:0 for the whole GPU session.

    No fallback: a missing libgpufem.so or a missing device is an error.
    """
    import gpufem_amd as gf
    ctx = gf.Context(0)
    yield ctx
    ctx.close()
