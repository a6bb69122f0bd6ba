import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtmgpu.so on the device)")


@pytest.fixture(scope="session", autouse=True)
def _built_libs():
    """Build the native libraries once if they are missing (CPU-side compile)."""
    from tmcore_amd import _ffi
    if not os.path.exists(os.path.join(_ffi.LIB_DIR, "libtmhost.so")):
        from tmcore_amd import build
        build.build_host()
    if not os.path.exists(os.path.join(_ffi.LIB_DIR, "libtmcpu.so")):
        from tmcore_amd import build
        build.build_fortran()
    oracle_so = os.path.join(ROOT, "oracle", "build", "liboracle.so")
    if not os.path.exists(oracle_so):
        from oracle import oracle
        oracle.build()
    yield
