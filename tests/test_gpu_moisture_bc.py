"""GPU parity under MoistureStateBC (RichardsModel): the library's hooks and
fused steps against the oracle.

Bars as in test_gpu_parity.py: hooks <= 1e-11 relative; a fused step <= 1e-10
relative on well-conditioned column-steps (the oracle's own output moves by
<= 1e-12 under its long-double and FMA probes), equal Newton counts on >= 98%
of the finite converged ones.
"""
import numpy as np
import pytest

from conftest import column_ok, rel_err

pytestmark = pytest.mark.gpu


def _batch(ncol=256, bottom=False):
    from clmsoil import inputs
    b = inputs.make_batch("rre_131072_uneven", ncol=ncol)
    return b.with_moisture_state(0.9, 0.6 if bottom else None, seed=3)


@pytest.mark.parametrize("bottom", [False, True])
def test_hooks_parity_moisture_state(gpu_lib, bottom):
    from clmsoil.model import (initialize_jacobian, make_compute_jacobian, make_update_cache,
                               make_update_implicit_cache, model_from_batch)
    from oracle.oracle import OracleModel
    b = _batch(bottom=bottom)
    m = model_from_batch(b, jac_precision=64)
    o = OracleModel.from_batch(b)
    tw = b.top_water[0].copy()
    bw = b.bottom_water.copy() if bottom else np.zeros(b.ncol)
    Y = b.copy_state()
    make_update_cache(m)(None, Y, 0.0)
    a = m.aux()
    oa = o.new_aux()
    o.update_cache(Y, oa, tw, None, bw)
    for k, ok in (("top_water", "top_water"), ("bottom_water", "bot_water")):
        e = rel_err(a[k], oa[ok], 1e-300)
        assert e.max() <= 1e-11, (k, e.max())
    assert np.abs(oa["top_water"]).max() > 0
    # cache_imp! at another state refreshes the boundary fluxes
    Y2 = {k: v.copy() for k, v in Y.items()}
    Y2["vartheta_l"] *= 1.01
    make_update_implicit_cache(m)(None, Y2, 0.0)
    o.update_implicit_cache(Y2, oa, tw, bw)
    a2 = m.aux()
    assert rel_err(a2["top_water"], oa["top_water"], 1e-300).max() <= 1e-11
    assert not np.allclose(a2["top_water"], a["top_water"])
    W = initialize_jacobian(m)
    make_compute_jacobian(m)(W, Y2, None, 1800.0, 0.0)
    ww, _, _ = m.get_jacobian()
    oww, _, _ = o.jacobian(Y2, oa, 1800.0)
    assert rel_err(ww, oww, 1e-300).max() <= 1e-11
    top_rows = 3 * (b.col_ptr[1:] - 1) + 1
    assert np.abs(oww[top_rows]).min() > 0
    m.close()


@pytest.mark.parametrize("solver", [pytest.param(1, id="group"), pytest.param(2, id="lane")])
@pytest.mark.parametrize("bottom", [False, True])
def test_fused_step_parity_moisture_state(gpu_lib, solver, bottom):
    from clmsoil.model import LandSimulation, model_from_batch
    from oracle.oracle import OracleModel
    b = _batch(ncol=384, bottom=bottom)
    w = b.workload
    m = model_from_batch(b, jac_precision=64, solver=solver)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    st, it = sim.solve(b.top_water[:1], None, 1, 1, want_iters=True)
    Yg, itg = sim.state(), it[0] & 0xFFFF
    tw = b.top_water[0].copy()
    bw = b.bottom_water.copy() if bottom else np.zeros(b.ncol)
    outs = {}
    for v in ("", "xprec", "fma"):
        o = OracleModel.from_batch(b, variant=v)
        Y = b.copy_state()
        ito, _ = o.step(Y, w.dt, w.newton_max_iters, w.newton_tol, tw, None, bw)
        outs[v] = (Y, ito)
    Yo, ito = outs[""]
    ok = column_ok(b, Yo) & column_ok(b, Yg) & (ito < w.newton_max_iters)
    sens = np.zeros(b.ncol)
    for v in ("xprec", "fma"):
        Yv, itv = outs[v]
        ok &= column_ok(b, Yv) & (itv == ito)
        for k in ("vartheta_l", "intF_water"):
            e = rel_err(Yv[k], Yo[k], 1e-12)
            if Yv[k].size == b.ncell:
                e = np.maximum.reduceat(np.where(np.repeat(ok, b.nlayers), e, 0.0),
                                        b.col_ptr[:-1])
            sens = np.maximum(sens, np.where(ok, e, 0.0))
    same = itg == ito
    print(f"moisture state (bottom={bottom}, solver={solver}): equal Newton counts "
          f"{same[ok].mean():.4f} on {ok.sum()} converged columns; mean iters "
          f"{itg.mean():.2f}", flush=True)
    assert ok.mean() > 0.5
    assert same[ok].mean() >= 0.98
    well = ok & same & (sens <= 1e-12)
    assert well.sum() >= 0.8 * ok.sum()
    for k in ("vartheta_l", "intF_water"):
        e = rel_err(Yg[k], Yo[k], 1e-12)
        if Yg[k].size == b.ncell:
            e = np.maximum.reduceat(np.where(np.repeat(well, b.nlayers), e, 0.0), b.col_ptr[:-1])
        e = np.where(well, e, 0.0)
        print(f"  {k}: max rel err {e.max():.2e} on {well.sum()} well-conditioned columns",
              flush=True)
        assert e.max() <= 1e-10, k
    m.close()
