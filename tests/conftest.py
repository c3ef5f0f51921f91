import importlib
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libnadapt.so")


@pytest.fixture(scope="session")
def na():
    return importlib.import_module("newton-adaptive_amd")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="session")
def nalib():
    L = importlib.import_module("newton-adaptive_amd._lib")
    return L
