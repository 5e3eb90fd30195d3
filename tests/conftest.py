import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def pytest_collection_modifyitems(config, items):
    """Tests marked slow run only with MCB_SLOW_TESTS=1 (they are full reference-size CPU runs)."""
    import pytest

    if os.environ.get("MCB_SLOW_TESTS") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU reference run; set MCB_SLOW_TESTS=1")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)
