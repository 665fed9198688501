import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def _gpu_available():
    # the AMD kernel driver's device node; present on a machine with a ROCm GPU
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    """GPU tests need a device; without one they are skipped here, while on a machine with a
    GPU a missing HIP library makes them fail (the library raises) rather than skip."""
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
