import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def load_package():
    """The product package (its directory name is not a Python identifier)."""
    return importlib.import_module("halide-sar-app_amd")


def load_scenes():
    return importlib.import_module("halide-sar-app_amd.scenes")


@pytest.fixture(scope="session")
def sar():
    return load_package()


@pytest.fixture(scope="session")
def scenes():
    return load_scenes()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
