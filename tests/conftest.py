import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    """The oracle with the library's own exp/log/log10/pow (the operations the GPU executes)."""
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def orc_libm():
    """The same oracle built with the C library's exp/log/log10/pow, i.e. the functions the
    reference Fortran calls: it reproduces the reference's golden outputs bit for bit."""
    from oracle.pyoracle import Oracle
    return Oracle(libm=True)


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name + ".npz"))
