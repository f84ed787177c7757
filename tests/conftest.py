import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcmtj_amd.so kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(autouse=True)
def _reset_constants():
    yield
    from cmtj_amd import constants

    constants.PhysicalConstants.resetToDefaults()
