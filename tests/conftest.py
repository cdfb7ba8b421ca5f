import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgat_hip.so on the device)")
    config.addinivalue_line("markers", "slow: full-size workload (minutes of CPU oracle time)")


@pytest.fixture(scope="session")
def gpu():
    """Initialises device 0 through the product library.  A GPU test fails
    (it is never skipped) when the native library or the device is missing."""
    import gat_amd
    gat_amd.init(0)
    return gat_amd


def pytest_collection_modifyitems(config, items):
    # Run GPU tests in file order but keep the slow full-size ones late and
    # the allocation-free entry point's test (tests/test_zz_no_alloc.py) last.
    def rank(it):
        if "test_zz_" in it.nodeid:
            return 2
        return 1 if it.get_closest_marker("slow") else 0

    items.sort(key=rank)
