import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the C-ABI")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    lib = ROOT / "oracle" / "liboracle.so"
    if not lib.exists():
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return dict(np.load(ROOT / "tests" / "golden" / "sparskit_ref.npz"))


@pytest.fixture(scope="session")
def gpu():
    """Initialise the device backend; the library must be built and loaded."""
    from getdp_amd import linalg
    linalg.init(0)
    return linalg
