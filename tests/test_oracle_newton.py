"""The oracle's exact EnergyHydrology Jacobian (ORC_JAC_NEWTON) against
central finite differences of the oracle's own residual.

R(x) = U0 + Δt T_imp(x) - x, where T_imp (energy_hydrology.jl:363-425) sees
x through the implicit cache (T and ψ, :427-445) and holds K, κ, θ_l and
the boundary fluxes at u_n.  Its derivative is the Newton matrix of the
mixed form's Newton option; the reference's modified-Picard matrix is a
different, lagged approximation, and the same check fails on it (a control
that the comparison can tell the two apart).  The Newton form converges to
the same root as the Picard form, in fewer iterations.
"""
import numpy as np
import pytest

from clmsoil import inputs


def _frozen_state(o, b, steps=6):
    """Advance a few Picard steps so that some layers hold ice."""
    w = b.workload
    Y = b.copy_state()
    for s in range(steps):
        sl = b.forcing_slot(s)
        o.step(Y, w.dt, w.newton_max_iters, w.newton_tol, b.top_water[sl].copy(),
               b.top_heat[sl].copy())
    return Y


def _residual(o, Yn, aux_n, x, dt):
    """Δt T_imp(x) - x with the cache of u_n and the implicit cache at x
    (U0 is a constant of the derivative)."""
    a = {k: v.copy() for k, v in aux_n.items()}
    o.update_implicit_cache(x, a)
    d = o.imp_tendency(x, a)
    return {k: dt * d[k] - x[k] for k in ("vartheta_l", "rho_e_int")}, a


def _fd_jacobian(o, b, Yn, x, dt):
    """Tridiagonal central differences: lower/diag/upper per cell for the
    blocks ϑϑ, ρeρe and ρeϑ (columns are independent, so perturbing every
    third cell of every column at once separates the entries)."""
    aux_n = o.new_aux()
    o.update_cache(Yn, aux_n, b.top_water[0].copy(), b.top_heat[0].copy())
    n = b.ncell
    cp = b.col_ptr
    lev = np.arange(n) - np.repeat(cp[:-1], np.diff(cp))
    W = {k: np.zeros(3 * n) for k in ("ww", "ee", "ew")}
    for var, hrel in (("vartheta_l", 1e-6), ("rho_e_int", 1e-7)):
        for color in range(3):
            sel = (lev % 3) == color
            h = np.zeros(n)
            h[sel] = hrel * np.maximum(np.abs(x[var][sel]), 1.0 if var == "rho_e_int" else 1e-2)
            xp = {k: v.copy() for k, v in x.items()}
            xm = {k: v.copy() for k, v in x.items()}
            xp[var] += h
            xm[var] -= h
            Rp, _ = _residual(o, Yn, aux_n, xp, dt)
            Rm, _ = _residual(o, Yn, aux_n, xm, dt)
            for rvar, blk in (("vartheta_l", "ww" if var == "vartheta_l" else None),
                              ("rho_e_int", "ew" if var == "vartheta_l" else "ee")):
                if blk is None:
                    continue
                dR = (Rp[rvar] - Rm[rvar])
                # entry (i, j) with j perturbed: i = j-1 (upper of row i), j, j+1 (lower)
                for off, slot in ((-1, 2), (0, 1), (1, 0)):
                    i = np.nonzero(sel)[0] + off
                    ok = (i >= 0) & (i < n)
                    j = np.nonzero(sel)[0][ok]
                    i = i[ok]
                    same_col = np.searchsorted(cp, i, side="right") == \
                        np.searchsorted(cp, j, side="right")
                    i, j = i[same_col], j[same_col]
                    W[blk][3 * i + slot] = dR[i] / (2.0 * h[j])
    return W, aux_n


def _kink_free(o, b, x, Yn):
    """Cells away from the non-smooth points of the residual: S = 1 (the
    saturated branch of ψ), ϑ_l = ν - θ_i (the cap of θ_l in T) and
    ϑ_l <= θ_r + √eps, below which effective_saturation clips ϑ_l
    (soil_hydrology_parameterizations.jl:45-50) and ψ stops depending on it."""
    P = b.params
    cp = b.col_ptr
    col = np.repeat(np.arange(b.ncol), np.diff(cp))
    nu, thr = P[0][col], P[1][col]
    vl = x["vartheta_l"]
    nue = nu - Yn["theta_i"]
    S = (vl - thr) / (nue - thr)
    return (np.abs(S - 1.0) > 1e-3) & (np.abs(vl - nue) > 1e-4) & (vl > thr + 1e-4)


def make_case(oracle_mod):
    b = inputs.make_batch("eh_262144_uneven", ncol=96, nsteps=12)
    o = oracle_mod.OracleModel.from_batch(b, jacobian=oracle_mod.JAC_NEWTON)
    Yn = _frozen_state(o, b)
    w = b.workload
    # the Newton iterate x: u_n plus part of one step's change
    x = {k: v.copy() for k, v in Yn.items()}
    Y1 = {k: v.copy() for k, v in Yn.items()}
    o.step(Y1, w.dt, w.newton_max_iters, w.newton_tol, b.top_water[0].copy(),
           b.top_heat[0].copy())
    for k in ("vartheta_l", "rho_e_int"):
        x[k] = Yn[k] + 0.6 * (Y1[k] - Yn[k])
    good = np.isfinite(x["vartheta_l"]) & np.isfinite(x["rho_e_int"])
    cols = [c for c in range(b.ncol) if good[b.col_ptr[c]:b.col_ptr[c + 1]].all()]
    sub = b.column_subset(np.array(cols))
    cell = np.concatenate([np.arange(b.col_ptr[c], b.col_ptr[c + 1]) for c in cols])
    colx = np.array(cols)
    Yn_s = {k: (v[cell] if v.size == b.ncell else v[colx]).copy() for k, v in Yn.items()}
    x_s = {k: (v[cell] if v.size == b.ncell else v[colx]).copy() for k, v in x.items()}
    o_s = oracle_mod.OracleModel.from_batch(sub, jacobian=oracle_mod.JAC_NEWTON)
    return oracle_mod, sub, o_s, Yn_s, x_s


@pytest.fixture(scope="module")
def case(oracle_mod):
    return make_case(oracle_mod)


def _analytic(oracle_mod, sub, Yn, x, jacobian):
    o = oracle_mod.OracleModel.from_batch(sub, jacobian=jacobian)
    aux = o.new_aux()
    o.update_cache(Yn, aux, sub.top_water[0].copy(), sub.top_heat[0].copy())
    o.update_implicit_cache(x, aux)
    ww, ee, ew = o.jacobian(x, aux, sub.workload.dt)
    return {"ww": ww, "ee": ee, "ew": ew}


def _row_rel(A, B, rows):
    """max over rows of |A - B| / max|row of B| (entries of one row share a scale)."""
    A3, B3 = A.reshape(-1, 3)[rows], B.reshape(-1, 3)[rows]
    scale = np.maximum(np.abs(B3).max(axis=1), 1e-300)
    return float((np.abs(A3 - B3).max(axis=1) / scale).max())


def test_newton_jacobian_matches_finite_differences(case):
    oracle_mod, sub, o, Yn, x = case
    assert (Yn["theta_i"] > 0).any(), "the case should hold some ice"
    fd, _ = _fd_jacobian(o, sub, Yn, x, sub.workload.dt)
    an = _analytic(oracle_mod, sub, Yn, x, oracle_mod.JAC_NEWTON)
    ok = _kink_free(o, sub, x, Yn)
    # a row is compared where it and its neighbours are away from the kinks
    rows = ok & np.roll(ok, 1) & np.roll(ok, -1)
    assert rows.mean() > 0.5
    for blk, tol in (("ww", 2e-5), ("ee", 2e-5), ("ew", 2e-4)):
        err = _row_rel(an[blk], fd[blk], rows)
        print(f"{blk}: max row-relative error vs finite differences {err:.2e}")
        assert err < tol, (blk, err)


def test_picard_jacobian_is_a_different_matrix(case):
    """The reference's modified-Picard blocks: the water block is the exact
    one; the energy blocks lack ∂T/∂ϑ and the advective ∂E/∂T, and A_ρϑ
    carries -I, so they fail the finite-difference check."""
    oracle_mod, sub, o, Yn, x = case
    fd, _ = _fd_jacobian(o, sub, Yn, x, sub.workload.dt)
    pic = _analytic(oracle_mod, sub, Yn, x, oracle_mod.JAC_PICARD)
    ok = _kink_free(o, sub, x, Yn)
    rows = ok & np.roll(ok, 1) & np.roll(ok, -1)
    assert _row_rel(pic["ww"], fd["ww"], rows) < 2e-5
    assert _row_rel(pic["ew"], fd["ew"], rows) > 1e-2
    assert _row_rel(pic["ee"], fd["ee"], rows) > 1e-3


def test_newton_form_same_root_fewer_iterations(oracle_mod):
    """Per-step map on 128 columns x 8 steps: from the same u_n, the Newton
    form's step agrees with the Picard form's to the Newton tolerance's
    reach where both converge, and takes fewer iterations."""
    b = inputs.make_batch("eh_262144_uneven", ncol=128, nsteps=8)
    w = b.workload
    op = oracle_mod.OracleModel.from_batch(b, jacobian=oracle_mod.JAC_PICARD)
    on = oracle_mod.OracleModel.from_batch(b, jacobian=oracle_mod.JAC_NEWTON)
    Y = b.copy_state()
    kp_all, kn_all, dv_all, de_all = [], [], [], []
    cp = b.col_ptr
    for s in range(8):
        sl = b.forcing_slot(s)
        Yp = {k: v.copy() for k, v in Y.items()}
        Yw = {k: v.copy() for k, v in Y.items()}
        kp, _ = op.step(Yp, w.dt, w.newton_max_iters, w.newton_tol, b.top_water[sl].copy(),
                        b.top_heat[sl].copy())
        kn, _ = on.step(Yw, w.dt, w.newton_max_iters, w.newton_tol, b.top_water[sl].copy(),
                        b.top_heat[sl].copy())
        fin = np.isfinite(Yp["vartheta_l"]) & np.isfinite(Yw["vartheta_l"])
        colfin = np.minimum.reduceat(fin.astype(int), cp[:-1]).astype(bool)
        both = colfin & (kp < w.newton_max_iters) & (kn < w.newton_max_iters)
        cells = np.repeat(both, np.diff(cp))
        rc = (np.repeat(b.params[12], np.diff(cp)) + 4.181e6 * np.minimum(Yp["vartheta_l"], 0.6))
        dv_all.append(np.abs(Yp["vartheta_l"] - Yw["vartheta_l"])[cells])
        de_all.append((np.abs(Yp["rho_e_int"] - Yw["rho_e_int"]) / rc)[cells])
        kp_all.append(kp[both])
        kn_all.append(kn[both])
        Y = Yp
    dv, de = np.concatenate(dv_all), np.concatenate(de_all)
    kp, kn = np.concatenate(kp_all), np.concatenate(kn_all)
    print(f"Picard mean iters {kp.mean():.2f}, Newton {kn.mean():.2f}; "
          f"max |dϑ| {dv.max():.2e}, max |dρe|/ρc {de.max():.2e} K")
    assert kn.mean() < 0.6 * kp.mean()
    assert dv.max() < 1e-6 and de.max() < 1e-6
