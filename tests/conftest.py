import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def golden():
    return np.load(ROOT / "tests" / "golden" / "statics.npz")


@pytest.fixture(scope="session")
def ctx():
    import randomtilings_amd as rt

    c = rt.Context(0)
    yield c
    c.close()


def square(a):
    a = np.asarray(a)
    n = int(round(np.sqrt(a.size)))
    assert n * n == a.size
    return a.reshape(n, n)
