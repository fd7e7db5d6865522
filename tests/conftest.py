import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name: str) -> dict:
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def golden_ham(g: dict) -> dict:
    """Hamiltonian dict (0-based abcd) from a golden fixture."""
    return dict(H0=complex(g["H0"]), H1=g["H1"], abcd=g["abcd"].astype(np.int64) - 1,
                H2=g["H2"].astype(np.float64), trperm=g["trperm"].astype(np.int64))


def golden_canon(g: dict) -> dict:
    return dict(D=g["zD0"].astype(complex), nocc=int(g["nocc0"]), nemp=int(g["nemp0"]),
                ovac=float(g["ovac0"]), voveru=g["voveru0"])


@pytest.fixture
def golden():
    return load_golden
