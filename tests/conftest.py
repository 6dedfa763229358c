import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libelegant_hip.so on cuda:0)")


@pytest.fixture(scope="session")
def tracker():
    from elegant_amd import native
    t = native.Tracker(0)
    yield t
    t.close()
