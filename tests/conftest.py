import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libm3stokes.so)")


@pytest.fixture(scope="session")
def c0():
    from m3tet_sph_amd import mesh as M
    return M.build_config("c0")


@pytest.fixture(scope="session")
def c0_inputs(c0):
    from m3tet_sph_amd import fields as F
    return F.make_inputs(c0, visc_contrast=1e3)


@pytest.fixture(scope="session")
def tiny():
    from m3tet_sph_amd import mesh as M
    return M.build_config("tiny")


def relerr(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
