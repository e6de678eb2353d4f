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


STALL_RESIDUAL = 1e-3  # the library's stall bound (ϑ_l; ρe_int/ρc_s in kelvin)


def oracle_exit_residual(o, batch, Yn, Yx, slot):
    """Per column, max |R(x)| of the step from Yn that ended at Yx, with the
    cache of u_n for the lagged fields and boundary fluxes and the implicit
    cache at x (R = U0 + Δt T_imp(x) - x; EnergyHydrology's ρe_int row scaled
    by ρc_s).  The library flags a column-step stalled when this residual of
    its last iteration exceeds STALL_RESIDUAL after the increments met tol."""
    w = batch.workload
    eh = batch.model == 1
    th = batch.top_heat[slot].copy() if eh else None
    a = o.new_aux()
    o.update_cache(Yn, a, batch.top_water[slot].copy(), th)
    U0 = {k: v.copy() for k, v in Yn.items()}
    if eh:
        e = o.exp_tendency(Yn, a)
        for k in U0:
            U0[k] = Yn[k] + w.dt * e[k]
    o.update_implicit_cache(Yx, a)
    d = o.imp_tendency(Yx, a)
    R = np.abs(U0["vartheta_l"] + w.dt * d["vartheta_l"] - Yx["vartheta_l"])
    if eh:
        rc = np.repeat(batch.params[12], batch.nlayers) + 4.181e6 * np.clip(Yx["vartheta_l"], 0, 1)
        R = np.maximum(R, np.abs(U0["rho_e_int"] + w.dt * d["rho_e_int"] - Yx["rho_e_int"]) / rc)
    return np.maximum.reduceat(np.where(np.isfinite(R), R, np.inf), batch.col_ptr[:-1])
