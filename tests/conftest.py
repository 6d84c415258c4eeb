import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _have_gpu():
    # Counting devices through the library does not need torch.
    try:
        from aces4_amd import sipgpu as sg
        return sg.device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if any(item.get_closest_marker("gpu") for item in items) and not _have_gpu():
        skip = pytest.mark.skip(reason="no GPU visible")
        for item in items:
            if item.get_closest_marker("gpu"):
                item.add_marker(skip)
