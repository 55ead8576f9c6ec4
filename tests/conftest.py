import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "acceleratedms-osm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def pytest_collection_modifyitems(config, items):
    # torch ships its own HIP runtime; in a process that uses both, torch's must start first (the
    # library's runtime starts lazily on its first HIP call). GPU sessions start it here.
    if any(item.get_closest_marker("gpu") for item in items):
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
