import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

MECH_DIR = os.path.join(ROOT, "hpc_ode_amd", "mechanisms")
GRI = os.path.join(MECH_DIR, "gri30.yaml")
H2 = os.path.join(MECH_DIR, "h2o2.yaml")
H2LI = os.path.join(MECH_DIR, "h2_li_2004.xml")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gri_path():
    return GRI


@pytest.fixture(scope="session")
def h2_path():
    return H2


@pytest.fixture(scope="session")
def h2li_path():
    return H2LI
