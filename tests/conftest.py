import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "cpm-rust_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running statistical test")
    config.addinivalue_line("markers", "exhaustive: runs for many minutes; only with "
                            "CPM_EXHAUSTIVE=1 in the environment")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("CPM_EXHAUSTIVE") == "1":
        return
    skip = pytest.mark.skip(reason="exhaustive check: set CPM_EXHAUSTIVE=1 to run it")
    for item in items:
        if "exhaustive" in item.keywords:
            item.add_marker(skip)
