import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_fixture(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def mps_arrays(z, prefix):
    n = int(z[prefix + "nsites"])
    a = [z[f"{prefix}a{i}"].copy() for i in range(n)]
    q = [z[f"{prefix}q{i}"].copy() for i in range(n + 1)]
    return z[prefix + "qsite"].copy(), q, a


def mpo_arrays(z, prefix="h_"):
    n = int(z[prefix + "nsites"])
    a = [z[f"{prefix}a{i}"].copy() for i in range(n)]
    q = [z[f"{prefix}q{i}"].copy() for i in range(n + 1)]
    return z[prefix + "qsite"].copy(), q, a


@pytest.fixture
def fixture_loader():
    return load_fixture
