import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def _ensure_built():
    """Build the native libraries once if they are missing (normally
    __graft_entry__.build() has built them)."""
    import subprocess
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    if not os.path.exists(os.path.join(ROOT, "freeshearlayer_amd", "lib", "libfsl.so")):
        from freeshearlayer_amd import build
        build.build_native()


def pytest_configure(config):
    _ensure_built()
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: longer CPU runs")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work so a fault is attributable
    if request.node.get_closest_marker("gpu") is not None:
        print("\n[gpu-case] %s" % request.node.nodeid, flush=True)
    yield
