import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# SALMON's testsuites/pseudo/Si_rps.dat (a data file of the reference's tests, KY format)
SI_RPS = os.path.join(ROOT, "tests", "data", "Si_rps.dat")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O

    if not os.path.exists(O.LIB):
        O.build()
    return O


@pytest.fixture(scope="session")
def tiny():
    from salmon2_amd import system

    return system.config("tiny")


@pytest.fixture(scope="session")
def c1():
    from salmon2_amd import system

    return system.config("c1")
