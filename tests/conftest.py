import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library's kernels)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def lib():
    from vare_amd import _lib as L
    L.load()
    return L
