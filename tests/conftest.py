import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (gfx950)")


def _has_gpu():
    try:
        from matops_amd import _lib
        return _lib.lib().mops_device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    # GPU tests run only where a GPU is visible; on a GPU box a missing
    # library is a failure, not a skip (the tests import it directly).
    if os.environ.get("MOPS_FORCE_GPU_TESTS"):
        return
    if not any("gpu" in it.keywords for it in items):
        return
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
