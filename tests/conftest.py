import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """GPU context; a GPU test fails (never skips) when the HIP library or the GPU is
    missing."""
    import viltrum_amd as V
    c = V.Context(0)
    yield c
    c.close()
