import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """Build libgdtp.so / liboracle.so only when they are missing (the tree ships them prebuilt)."""
    from gradientdomaintextureprocessing_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from gradientdomaintextureprocessing_amd import build
        build.build()
    import oracle
    oracle.load()
    yield
