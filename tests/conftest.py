import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libchakum.so")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(GOLDEN / name, allow_pickle=False)
    return load


@pytest.fixture(scope="session")
def ctx():
    from chakumrender_amd import abi

    c = abi.Context(0)  # raises ChkError when no HIP device is visible
    yield c
    c.close()


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)
