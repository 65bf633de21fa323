import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


def pytest_collection_modifyitems(config, items):
    # keep CPU runs from touching the GPU library's compute paths by accident
    pass


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
