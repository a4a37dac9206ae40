import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dosecuda-jax_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

MACHINE_DIR = os.path.join(PKG, "DoseCUDA", "lookuptables", "protons", "HitachiProbeatJHU")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X / gfx950)")


@pytest.fixture(scope="session")
def machine_dir():
    return MACHINE_DIR


@pytest.fixture(scope="session")
def plan_models():
    from DoseCUDA import IMPTPlan
    return IMPTPlan()
