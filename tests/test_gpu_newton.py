"""GPU: the Newton form of the EnergyHydrology step (CLM_JACOBIAN_NEWTON)
against the FP64 oracle.

The Newton form solves the same FP64 residual to the same stopping rule with
the exact Jacobian (the oracle's ORC_JAC_NEWTON, pinned by finite differences
in test_oracle_newton.py), so it converges to the reference's root in fewer
iterations.  Bars:
  * against the reference-Jacobian FP64 oracle (the reference's own method):
    |Δϑ_l| <= 1e-6 and |Δρe_int|/ρc_s <= 1e-6 K per step on column-steps that
    both converge and stay finite;
  * FP64 Newton form against the oracle's Newton form: 1e-10 relative on
    well-conditioned column-steps with equal Newton counts, equal counts on
    >= 97% of them;
  * mixed (FP32 matrix) Newton form: mean Newton count within 0.5 of the
    oracle's Newton form and below 0.6 x the reference method's.
"""
import numpy as np
import pytest

from conftest import STALL_RESIDUAL, column_ok, oracle_exit_residual, rel_err

pytestmark = pytest.mark.gpu

MENU = [(8, 1), (16, 1), (32, 1), (64, 1), (8, 2), (16, 2), (32, 2), (64, 2), (8, 3), (16, 3),
        (32, 3), (4, 4), (8, 4), (16, 4), (32, 4), (8, 5), (16, 5), (8, 6), (8, 8), (16, 8)]


def _rhoc(b, Y):
    """ρc_s with θ_l = ϑ_l (the norm's scale; the exact θ_l cap is immaterial here)."""
    return np.repeat(b.params[12], b.nlayers) + 4.181e6 * np.clip(Y["vartheta_l"], 0, 1)


def _newton_model(b, jac):
    from clmsoil import _lib as L
    from clmsoil.model import model_from_batch
    return model_from_batch(b, jac_precision=jac, solver=L.SOLVER_GROUP,
                            jacobian=L.JACOBIAN_NEWTON)


def _oracle_step(b, jacobian, Y):
    from oracle.oracle import OracleModel
    w = b.workload
    o = OracleModel.from_batch(b, jacobian=jacobian)
    Y = {k: v.copy() for k, v in Y.items()}
    it, _ = o.step(Y, w.dt, w.newton_max_iters, w.newton_tol, b.top_water[0].copy(),
                   b.top_heat[0].copy())
    return Y, it


@pytest.mark.parametrize("jac", [32, 64])
def test_newton_form_every_shape(gpu_lib, jac):
    """Every compiled shape that fits, one converged step of the Newton form:
    against the reference method's FP64 result at 1e-6, and (FP64 Newton
    form) against the oracle's Newton form at 1e-10 with equal counts."""
    from clmsoil import inputs
    from clmsoil.model import LandSimulation
    from oracle.oracle import JAC_NEWTON, JAC_PICARD
    b = inputs.make_batch("eh_262144_uneven", ncol=384)
    w = b.workload
    Yp, itp = _oracle_step(b, JAC_PICARD, b.copy_state())
    Yn, itn = _oracle_step(b, JAC_NEWTON, b.copy_state())
    from oracle.oracle import OracleModel
    stall = oracle_exit_residual(OracleModel.from_batch(b), b, b.copy_state(), Yp, 0) > \
        STALL_RESIDUAL
    m = _newton_model(b, jac)
    buckets = [(mx, n) for (mx, G, R, n) in m.shapes() if n > 0]
    nmax = {mx: int(b.nlayers[(b.nlayers <= mx) & (b.nlayers > prev)].max())
            for (mx, _), prev in zip(buckets, [0] + [x for x, _ in buckets[:-1]])}
    rc = _rhoc(b, Yp)
    tried = 0
    for G, R in MENU:
        fits = [mx for mx, _ in buckets if nmax[mx] <= G * R]
        if not fits:
            continue
        for mx in fits:
            m.set_shape(mx, G, R)
        sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                             tol=w.newton_tol)
        _, it = sim.solve(b.top_water[:1], b.top_heat[:1], 1, 1, want_iters=True)
        it = it[0]
        Yg = sim.state()
        in_fit = np.isin(np.searchsorted([mx for mx, _ in buckets], b.nlayers),
                         [i for i, (mx, _) in enumerate(buckets) if mx in fits])
        conv = ((it & 0x30000) == 0) & (itp < w.newton_max_iters) & column_ok(b, Yg) & \
            column_ok(b, Yp) & ~stall
        cm = np.repeat(in_fit & conv, b.nlayers)
        dv = np.abs(Yg["vartheta_l"][cm] - Yp["vartheta_l"][cm]).max()
        de = (np.abs(Yg["rho_e_int"][cm] - Yp["rho_e_int"][cm]) / rc[cm]).max()
        kg = (it & 0xFFFF)[in_fit & conv]
        print(f"newton jac {jac} shape ({G}, {R}) on {fits}: vs reference method |dϑ| {dv:.2g}, "
              f"|dρe|/ρc {de:.2g} K; mean iters {kg.mean():.2f} (reference method "
              f"{itp[in_fit & conv].mean():.2f}, oracle Newton {itn[in_fit & conv].mean():.2f})",
              flush=True)
        assert cm.sum() > 0 and dv <= 1e-6 and de <= 1e-6, (G, R, dv, de)
        if jac == 64:
            same = in_fit & conv & ((it & 0xFFFF) == itn) & column_ok(b, Yn)
            assert same.sum() >= 0.97 * (in_fit & conv).sum(), (G, R)
            for k, fl in (("vartheta_l", 1e-12), ("rho_e_int", 1.0)):
                e = rel_err(Yg[k], Yn[k], fl)[np.repeat(same, b.nlayers)]
                assert e.max() <= 1e-10, (G, R, k, e.max())
        tried += 1
    assert tried >= 6
    m.close()


def test_newton_form_per_step_map(gpu_lib):
    """12 steps of 512 columns, each step from the reference oracle's state:
    the mixed Newton form against the reference method (FP64 oracle) and
    against the oracle's Newton form."""
    from clmsoil import inputs
    from clmsoil.model import LandSimulation
    from oracle.oracle import JAC_NEWTON, JAC_PICARD, OracleModel
    b = inputs.make_batch("eh_262144_uneven", ncol=512, nsteps=12)
    w = b.workload
    m = _newton_model(b, 32)
    op = OracleModel.from_batch(b, jacobian=JAC_PICARD)
    on = OracleModel.from_batch(b, jacobian=JAC_NEWTON)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                         tol=w.newton_tol)
    Y = b.copy_state()
    held = total = nfb = nstall = 0
    wv = we = 0.0
    kg_all, kp_all, kn_all = [], [], []
    for s in range(12):
        sl = b.forcing_slot(s)
        sim.reset(Y)
        _, it = sim.solve(b.top_water[sl:sl + 1], b.top_heat[sl:sl + 1], 1, 1, want_iters=True)
        it = it[0]
        Yg = sim.state()
        Yp = {k: v.copy() for k, v in Y.items()}
        Yn = {k: v.copy() for k, v in Y.items()}
        fw, fh = b.top_water[sl].copy(), b.top_heat[sl].copy()
        itp, _ = op.step(Yp, w.dt, w.newton_max_iters, w.newton_tol, fw, fh)
        itn, _ = on.step(Yn, w.dt, w.newton_max_iters, w.newton_tol, fw, fh)
        ok = column_ok(b, Y) & column_ok(b, Yg) & column_ok(b, Yp) & column_ok(b, Yn)
        # column-steps where the reference method stops away from a root
        # (its increments met tol with a large residual) are no reference
        stall = oracle_exit_residual(op, b, Y, Yp, sl) > STALL_RESIDUAL
        nstall += (stall & (itp < w.newton_max_iters)).sum()
        conv = ok & ((it & 0x30000) == 0) & (itp < w.newton_max_iters) & \
            (itn < w.newton_max_iters) & ~stall
        cm = np.repeat(conv, b.nlayers)
        rc = _rhoc(b, Yp)
        dv = np.abs(Yg["vartheta_l"] - Yp["vartheta_l"])
        de = np.abs(Yg["rho_e_int"] - Yp["rho_e_int"]) / rc
        bad = np.maximum.reduceat(np.where(cm, np.maximum(dv, de), 0.0), b.col_ptr[:-1]) > 1e-6
        held += (conv & ~bad).sum()
        total += conv.sum()
        wv = max(wv, dv[cm].max())
        we = max(we, de[cm].max())
        nfb += ((it & 0x80000) != 0).sum()
        kg_all.append((it & 0xFFFF)[conv])
        kp_all.append(itp[conv])
        kn_all.append(itn[conv])
        Y = Yp
    kg, kp, kn = (np.concatenate(a) for a in (kg_all, kp_all, kn_all))
    print(f"newton mixed per-step map: {held} of {total} converged column-steps within 1e-6 "
          f"of the reference method (max |dϑ| {wv:.2e}, |dρe|/ρc {we:.2e} K); mean Newton "
          f"iters GPU {kg.mean():.2f}, oracle Newton {kn.mean():.2f}, reference method "
          f"{kp.mean():.2f}; {nfb} column-steps fell back; the reference method stalled on "
          f"{nstall}", flush=True)
    assert held == total and total > 0.9 * 12 * 512
    assert kg.mean() <= kn.mean() + 0.5
    assert kg.mean() <= 0.6 * kp.mean()


def test_newton_form_rejects_unsupported(gpu_lib):
    from clmsoil import _lib as L
    from clmsoil import inputs
    from clmsoil._lib import ClmError
    from clmsoil.model import model_from_batch
    with pytest.raises(ClmError):  # RichardsModel has no Newton form here
        model_from_batch(inputs.make_batch("rre_131072_uneven", ncol=32),
                         jacobian=L.JACOBIAN_NEWTON)
    with pytest.raises(ClmError):  # the lane solver neither
        model_from_batch(inputs.make_batch("eh_262144_uneven", ncol=32),
                         solver=L.SOLVER_LANE, jacobian=L.JACOBIAN_NEWTON)
