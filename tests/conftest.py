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
    """Build libgdtp.so / liboracle.so when they are missing or libgdtp.so was built from other
    sources than those in the tree (the tree normally ships them prebuilt and current)."""
    from gradientdomaintextureprocessing_amd import _lib
    if not _lib.is_current():
        from gradientdomaintextureprocessing_amd import build
        build.build()
    import oracle
    oracle.load()
    yield
