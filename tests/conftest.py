import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs on the MI355X box)")


def have_reference_harness():
    return os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libsemba_ref.so"))


needs_ref = pytest.mark.skipif(not have_reference_harness(),
                               reason="reference harness oracle/_ref not built here")
