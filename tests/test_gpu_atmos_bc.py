"""GPU parity of EnergyHydrology's AtmosDrivenFluxBC (soil-only form, supplied
surface fluxes) with NoRunoff and TOPMODELRunoff: the library's hooks and
fused group step against the oracle.

Bars as in test_gpu_parity.py: hooks <= 1e-11 relative (tendencies against
the column's scale); a fused step <= 1e-10 relative on well-conditioned
column-steps, with equal Newton counts on >= 98% of the converged ones.
The saturated heights are summed in another order on the GPU (a group
reduction), which the 1e-11 hook bar covers.
"""
import numpy as np
import pytest

from conftest import column_ok, rel_err

pytestmark = pytest.mark.gpu

KINDS = {"none": False, "topmodel": True}


def _batch(kind, ncol=256):
    from clmsoil import inputs
    return inputs.make_batch("eh_262144_uneven", ncol=ncol).with_atmos(KINDS[kind], seed=9)


def _sfc(b, slot=0):
    return {k: v[slot].copy() for k, v in b.surface.items()}


@pytest.mark.parametrize("kind", list(KINDS))
def test_hooks_parity_atmos(gpu_lib, kind):
    from clmsoil.model import (make_compute_imp_tendency, make_exp_tendency,
                               make_update_cache, model_from_batch)
    from oracle.oracle import OracleModel
    b = _batch(kind)
    m = model_from_batch(b, jac_precision=64, solver=1)
    o = OracleModel.from_batch(b)
    Y = b.copy_state()
    make_update_cache(m)(None, Y, 0.0)
    a = m.aux()
    oa = o.new_aux()
    o.update_cache(Y, oa, None, None, None, None, sfc=_sfc(b))
    keys = [("top_water", "top_water"), ("top_heat", "top_heat"), ("bottom_water", "bot_water")]
    if KINDS[kind]:
        keys += [("sink", "sink"), ("r_ss", "r_ss"), ("sink_energy", "sink_e"),
                 ("r_ess", "r_ess")]
        assert (oa["sink"] > 0).any() and (oa["r_ss"] > 0).all()
    for k, ok in keys:
        e = rel_err(a[k], oa[ok], 1e-300)
        print(f"{kind} {k}: max rel err {e.max():.2e}", flush=True)
        assert e.max() <= 1e-11, (k, e.max())
    dE = {k: np.zeros_like(v) for k, v in Y.items()}
    make_exp_tendency(m)(dE, Y, None, 0.0)
    odE = o.exp_tendency(Y, oa)
    assert (odE["theta_i"] != 0).any()
    for k in ("vartheta_l", "theta_i", "intF_water"):
        scale = np.abs(odE[k]).max()
        assert (np.abs(dE[k] - odE[k]) <= 1e-11 * scale + 1e-300).all(), k
    dY = {k: np.zeros_like(v) for k, v in Y.items()}
    make_compute_imp_tendency(m)(dY, Y, None, 0.0)
    od = o.imp_tendency(Y, oa)
    for k in ("vartheta_l", "rho_e_int"):
        scale = np.repeat(np.maximum.reduceat(np.abs(od[k]), b.col_ptr[:-1]), b.nlayers)
        assert (np.abs(dY[k] - od[k]) <= 1e-10 * scale + 1e-300).all(), k
    for k in ("intF_water", "intF_energy"):
        assert rel_err(dY[k], od[k], 1e-300).max() <= 1e-12, k
    m.close()


@pytest.mark.parametrize("kind", list(KINDS))
def test_fused_step_parity_atmos(gpu_lib, kind):
    from clmsoil.model import LandSimulation, model_from_batch
    from oracle.oracle import OracleModel
    b = _batch(kind, ncol=384)
    w = b.workload
    m = model_from_batch(b, jac_precision=64, solver=1)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    st, it = sim.solve(None, None, 1, 1, want_iters=True,
                       surface={k: v[:1] for k, v in b.surface.items()})
    Yg, itg = sim.state(), it[0] & 0xFFFF
    outs = {}
    for v in ("", "xprec", "fma"):
        o = OracleModel.from_batch(b, variant=v)
        Y = b.copy_state()
        ito, _ = o.step(Y, w.dt, w.newton_max_iters, w.newton_tol, None, None, None, None,
                        sfc=_sfc(b))
        outs[v] = (Y, ito)
    Yo, ito = outs[""]
    keys = ("vartheta_l", "theta_i", "rho_e_int", "intF_water", "intF_energy")
    floors = {"vartheta_l": 1e-12, "theta_i": 1e-3, "rho_e_int": 1.0, "intF_water": 1e-12,
              "intF_energy": 1.0}
    ok = column_ok(b, Yo) & column_ok(b, Yg) & (ito < w.newton_max_iters)
    sens = np.zeros(b.ncol)
    for v in ("xprec", "fma"):
        Yv, itv = outs[v]
        ok &= column_ok(b, Yv) & (itv == ito)
        for k in keys:
            e = rel_err(Yv[k], Yo[k], floors[k])
            if Yv[k].size == b.ncell:
                e = np.maximum.reduceat(np.where(np.repeat(ok, b.nlayers), e, 0.0),
                                        b.col_ptr[:-1])
            sens = np.maximum(sens, np.where(ok, e, 0.0))
    same = itg == ito
    print(f"atmos {kind}: equal Newton counts {same[ok].mean():.4f} on {ok.sum()} converged "
          f"columns; mean iters {itg.mean():.2f}", flush=True)
    assert ok.mean() > 0.5
    assert same[ok].mean() >= 0.98
    well = ok & same & (sens <= 1e-12)
    assert well.sum() >= 0.8 * ok.sum()
    for k in keys:
        e = rel_err(Yg[k], Yo[k], floors[k])
        if Yg[k].size == b.ncell:
            e = np.maximum.reduceat(np.where(np.repeat(well, b.nlayers), e, 0.0), b.col_ptr[:-1])
        e = np.where(well, e, 0.0)
        print(f"  {k}: max rel err {e.max():.2e} on {well.sum()} well-conditioned columns",
              flush=True)
        assert e.max() <= 1e-10, k
    # sublimation reached θ_i: the columns below freezing lost ice to the air
    assert (b.surface["vapor_flux_ice"][0] > 0).any()
    m.close()


def test_run_with_surface_table_equals_steps(gpu_lib):
    """clm_run over a two-slot surface-flux table equals step-by-step runs."""
    from clmsoil.model import LandSimulation, model_from_batch
    b = _batch("topmodel", ncol=200)
    w = b.workload
    tab = {k: v[:2] for k, v in b.surface.items()}
    m = model_from_batch(b, jac_precision=64, solver=1)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    sim.solve(None, None, 4, 2, surface=tab)
    Yr = sim.state()
    sim.reset(b.copy_state())
    for s in range(4):
        sl = s // 2
        sim.solve(None, None, 1, 1, surface={k: v[sl:sl + 1] for k, v in tab.items()})
    Ys = sim.state()
    for k in Yr:
        assert np.array_equal(Yr[k], Ys[k], equal_nan=True), k
    m.close()
