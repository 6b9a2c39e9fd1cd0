import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for _p in (ROOT, ROOT / "tests"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libgeoswe_hip.so")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is visible to the HIP library (the library itself must load)."""
    from geoswe_amd import _lib
    _lib.lib()     # raises ImportError if the library was not built: a loud failure, not a skip
    if _lib.device_count() < 1:
        pytest.skip("no GPU visible")
    return True
