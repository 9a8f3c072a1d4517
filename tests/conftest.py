import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU relaxations")


@pytest.fixture(scope="session")
def W():
    return importlib.import_module("weave-sim_amd")


@pytest.fixture(scope="session")
def gpu_ctx(W):
    """the HIP context; GPU tests fail loudly (not skip) when the device library cannot run"""
    return W.solver.default_context()
