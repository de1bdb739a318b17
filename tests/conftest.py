import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built libann_amd.so")


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.lib()
    return pyoracle
