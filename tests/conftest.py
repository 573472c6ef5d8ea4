import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built libtmce.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


SETS = ["348864", "460896", "6688128", "6960119", "8192128"]


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle
