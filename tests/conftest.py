import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, loaded; fails loudly if it is missing or no device is visible."""
    from aniso_amd import _lib

    L = _lib.load()
    assert L.ans_available() == 1, "no HIP device visible"
    return L
