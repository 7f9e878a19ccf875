"""Test configuration: registers the ``gpu`` marker and skips GPU tests only when no AMD
GPU device node exists.  On a machine with a GPU, a missing or broken ``libgsw_hip.so``
makes the GPU tests fail instead of skipping."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HAS_GPU_NODE = os.path.exists("/dev/kfd")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libgsw_hip.so")


def pytest_collection_modifyitems(config, items):
    if HAS_GPU_NODE:
        return
    skip = pytest.mark.skip(reason="no AMD GPU device node (/dev/kfd)")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
