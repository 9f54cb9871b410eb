import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libgls kernels)")


@pytest.fixture(scope="session")
def gls_lib():
    """libgls.so loaded (built in-tree if missing); no device calls."""
    from cuvarbase_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        from cuvarbase_amd import build

        build.build()
    return _lib.load()
