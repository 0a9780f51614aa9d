import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfe_hip.so)")


@pytest.fixture(scope="session")
def oracle():
    from tests.oracle_bind import load_oracle
    return load_oracle()


def golden(name):
    import numpy as np
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))
