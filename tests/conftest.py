import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "climaland.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library on a machine with a GPU: fail loudly if it is missing."""
    import torch  # noqa: F401  (device presence check only)
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from clmsoil import _lib
    lib = _lib.load()
    assert os.path.exists(_lib.LIB_PATH)
    return lib


def rel_err(a, b, floor=1e-12):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def column_ok(batch, Y, vmax=5.0):
    """Columns whose state stayed finite and bounded (|ϑ_l| < vmax)."""
    vl = Y["vartheta_l"]
    ok = np.isfinite(vl) & (np.abs(vl) < vmax)
    if "rho_e_int" in Y and Y.get("rho_e_int") is not None:
        ok &= np.isfinite(Y["rho_e_int"]) & np.isfinite(Y["theta_i"])
    return np.minimum.reduceat(ok.astype(np.int8), batch.col_ptr[:-1]).astype(bool)
