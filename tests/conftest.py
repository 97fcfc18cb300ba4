import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import mbs_loader  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    return mbs_loader.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    return o


def load_fixture(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


@pytest.fixture(scope="session")
def fixtures():
    return {n: load_fixture(n) for n in ("config1_fixed", "fixed_absorbing", "sobol_average",
                                         "config5_near_forward")}
