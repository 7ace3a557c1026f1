import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: large sizes")


def pytest_collection_modifyitems(config, items):
    # GPU tests first inside each file is not needed; keep file order.
    return


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref_mod(oracle_mod):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref/libref.so not built (needs the reference sources)")
    oracle_mod.ref()
    return oracle_mod


@pytest.fixture(scope="session")
def fp():
    """The product package with its HIP library loaded (GPU tests only)."""
    import fft_method_test_amd as F
    from fft_method_test_amd import _lib

    _lib.load()
    return F


def random_mask(shape, porosity, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    return (rng.random(nx * ny * nz) < porosity).astype(np.uint8)


def rel(a, b):
    """Relative Frobenius distance ||a - b|| / ||b|| (real or complex)."""
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d
