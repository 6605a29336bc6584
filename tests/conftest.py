import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

# The reference library (built only in the build container) must be loaded before large
# numpy work; its static heap set-up crashes when dlopen'ed later in the process.
try:
    from oracle import harness as _h
    if _h.ref_available() and os.path.isdir("/root/reference/Source"):
        _h.load_ref()
except Exception:  # the reference build is optional
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "reference: needs the reference renderer built from source in oracle/_ref")
