import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libmlh_hip.so")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP library; a GPU test fails loudly when it is missing."""
    from meshlesshydro_amd import lib
    L = lib.lib()
    import ctypes
    n = ctypes.c_int(0)
    return lib
