import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def angles(a, b):
    """Angle between row vectors, robust near 0: atan2(|a x b|, a . b) in fp64."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.sum(a * b, 1))


def axis_angles(a, b):
    """Angle between the lines spanned by the rows (sign ignored)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.abs(np.sum(a * b, 1)))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle
