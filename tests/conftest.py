import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "incremental-fluids_amd")
for p in (os.path.join(ROOT, "tests"), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libifl.so")


def gpu_available():
    try:
        import ifl_fluids
        return ifl_fluids.lib is not None
    except Exception:
        return False
