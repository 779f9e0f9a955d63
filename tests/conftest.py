import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def hiplib():
    """The built C-ABI library; builds it only if nothing is built yet."""
    from gpusgdlm_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        from gpusgdlm_amd import build
        build.build(verbose=False)
    return _native.load()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    return oracle.lib()
