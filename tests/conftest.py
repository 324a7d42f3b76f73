import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def bunny():
    return dict(np.load(os.path.join(GOLDEN, "bunny.npz")))


@pytest.fixture(scope="session")
def elements_golden():
    return dict(np.load(os.path.join(GOLDEN, "elements.npz")))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.load()
    return O
