import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

EPS = np.finfo(np.float64).eps
# MAGMA testers accept errors below tolerance * eps with tolerance = 30
# (testing/magma_util.cpp:138,192).
TESTER_TOL = 30.0 * EPS


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")


def uniform_batch(rng, batch, n, ld=None, cols=None):
    """(batch, cols, ld) column-major batch with U(-1,1) entries in the
    leading n rows and zeros in the padding."""
    ld = n if ld is None else ld
    cols = n if cols is None else cols
    A = np.zeros((batch, cols, ld))
    A[:, :, :n] = rng.uniform(-1.0, 1.0, (batch, cols, n))
    return A


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc.lib()
    return orc


@pytest.fixture(scope="session")
def golden():
    return np.load(ROOT / "tests" / "golden" / "lu_fixtures.npz")


def golden_cases(g):
    return sorted({k.split("/")[0] for k in g.files if not k.startswith("_")})
