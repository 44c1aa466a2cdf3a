"""Shared test setup: paths, the `gpu` marker, and small model builders."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "topo-dda_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)

# objective of python_tests/test_DDAModelWrapper.py:40 (22x22x10, FCD, +-50 images)
GOLDEN_OBJECTIVE = 11.005338768207627


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def tables():
    import dda_oracle as O
    return O.sici_tables()


@pytest.fixture(scope="session")
def golden_inputs():
    diel = np.loadtxt(os.path.join(GOLDEN, "dielectrics.txt"))
    geo = np.loadtxt(os.path.join(GOLDEN, "geometry.txt"), dtype=int)
    return diel, geo
