import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

OPTICS_DIR = os.path.join(ROOT, "tests", "golden", "lattices")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP device)")


def pytest_sessionstart(session):
    """Build libufo_hip.so once if nothing is built yet (a fresh checkout);
    a library already in the tree is used as it is."""
    from ufo_amd import build
    if not os.path.exists(build.LIB):
        build.build_library()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no HIP device in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
