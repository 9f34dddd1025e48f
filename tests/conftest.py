import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(GOLDEN, "streams.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def graph_path():
    def path(name):
        return os.path.join(GOLDEN, "graphs", name + ".mtx")
    return path
