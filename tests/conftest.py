import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def pytest_collection_modifyitems(config, items):
    # deterministic order: CPU tests first is not required, keep file order
    return


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def nthreads():
    # at most 16 oracle threads (the CPU share next to one GPU)
    return max(1, min(16, os.cpu_count() or 1))
