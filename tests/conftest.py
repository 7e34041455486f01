import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _built(path):
    return os.path.exists(path)


@pytest.fixture(scope="session", autouse=True)
def _ensure_built():
    """Tests load what is built; they build only when nothing is."""
    from mnhack8_amd import build
    if not _built(build.LIB) or not _built(build.DRIVER):
        build.build_library()
        build.build_driver()
    if not _built(os.path.join(ROOT, "oracle", "_build", "libheat_oracle.so")):
        build.build_oracle()
