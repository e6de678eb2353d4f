"""Oracle column operators: the reference's Jacobian KATs
(test/shared_utilities/implicit_timestepping/{richards,energy_hydrology}_model.jl),
the hydrostatic KAT (test/standalone/Soil/soiltest.jl:15-90), and
independent checks: finite-difference Jacobians, dense solves, discrete
conservation, and the converged step as the root of the implicit-Euler
residual found by a general-purpose dense Newton (scipy)."""
import numpy as np
import pytest

from conftest import rel_err


def uniform_column(zmin, zmax, n):
    return np.linspace(zmin, zmax, n + 1)


def single_column(oracle_mod, model, bc_bottom, zf, params):
    N = zf.size - 1
    cp = np.array([0, N], dtype=np.int32)
    P = np.asarray(params, dtype=np.float64).reshape(-1, 1)
    return oracle_mod.OracleModel(model, bc_bottom, cp, zf, P)


RRE_KAT = dict(nu=0.495, Ksat=0.0443 / 3600 / 100, Ss=1e-3, n=1.43, a=2.6, thr=0.124)


def rre_params(p):
    return [p["nu"], p["thr"], p["a"], p["n"], p["Ksat"], p["Ss"]]


def test_rre_jacobian_kat_flux_bc(oracle_mod):
    """richards_model.jl:147-227: uniform 150-layer column, ϑ = 0.24, dtγ = 1."""
    p = RRE_KAT
    o = single_column(oracle_mod, 0, 0, uniform_column(-1.5, 0.0, 150), rre_params(p))
    Y = o.new_state()
    Y["vartheta_l"][:] = 0.24
    aux = o.new_aux()
    o.update_cache(Y, aux, np.array([-p["Ksat"]]), None, np.array([0.0]))
    W, _, _ = o.jacobian(Y, aux, 1.0)
    W = W.reshape(-1, 3)
    L = oracle_mod.lib()
    K = L.orc_hydraulic_conductivity(p["n"], p["Ksat"],
                                     L.orc_effective_saturation(p["nu"], 0.24, p["thr"]))
    D = L.orc_dpsidtheta(p["a"], p["n"], 0.24, p["nu"], p["thr"], p["Ss"])
    dz = 0.01
    assert W[0, 0] == 0.0 and W[-1, 2] == 0.0
    assert np.allclose(W[1:, 0], W[:-1, 2], rtol=1e-12)
    assert np.allclose(W[1:, 0], K / dz ** 2 * D, rtol=1e-9)
    assert W[0, 1] == pytest.approx(-K / dz ** 2 * D - 1, rel=1e-12)
    assert np.allclose(W[1:-1, 1], -2 * K / dz ** 2 * D - 1, rtol=1e-12)
    assert W[-1, 1] == pytest.approx(-K / dz ** 2 * D - 1, rel=1e-12)


def test_eh_jacobian_kat(oracle_mod):
    """energy_hydrology_model.jl:16-174: ϑ = 0.24, θ_i = 0, T = 280 K, zero
    flux BCs, quartz-only solids; checks all three blocks incl. the -I on
    the diagonal of ∂ρe/∂ϑ."""
    L = oracle_mod.lib()
    K_ = oracle_mod.constants()
    p = RRE_KAT
    nu, om, q, gr = p["nu"], 0.0, 1.0, 0.0
    ks = L.orc_kappa_solid(om, q, 0.25, 8.0, 2.5)
    rp = q * 2.66e3 + (1 - om - q - gr) * 2.65e3
    k_dry = L.orc_kappa_dry(rp, nu, ks, K_["K_THERM"])
    k_f = L.orc_kappa_sat_frozen(ks, nu, 2.21)
    k_u = L.orc_kappa_sat_unfrozen(ks, nu, 0.57)
    rc_ds = (1 - nu) * (q * 2.01e6 + (1 - q) * 2.01e6)
    params = rre_params(p) + [om, q, gr, k_dry, k_f, k_u, rc_ds]
    o = single_column(oracle_mod, 1, 0, uniform_column(-1.5, 0.0, 150), params)
    Y = o.new_state()
    Y["vartheta_l"][:] = 0.24
    rcs = L.orc_volumetric_heat_capacity(0.24, 0.0, rc_ds)
    Y["rho_e_int"][:] = L.orc_volumetric_internal_energy(0.0, rcs, 280.0)
    aux = o.new_aux()
    z = np.zeros(1)
    o.update_cache(Y, aux, z, z, z, z)
    Ww, We, Wew = (w.reshape(-1, 3) for w in o.jacobian(Y, aux, 1.0))
    K = L.orc_impedance_factor(0.0, 7.0) * L.orc_viscosity_factor(280.0, 2.64e-2, 288.0) * \
        L.orc_hydraulic_conductivity(p["n"], p["Ksat"], L.orc_effective_saturation(nu, 0.24, p["thr"]))
    D = L.orc_dpsidtheta(p["a"], p["n"], 0.24, nu, p["thr"], p["Ss"])
    kap = L.orc_thermal_conductivity(
        k_dry, L.orc_kersten_number(0.0, L.orc_relative_saturation(0.24, 0.0, nu), 0.24, 18.3,
                                    om, q, gr), L.orc_kappa_sat(0.24, 0.0, k_u, k_f))
    dTdr = 1.0 / rcs
    el = L.orc_volumetric_internal_energy_liq(280.0)
    dz = 0.01
    rt = 1e-9  # T recovered from ρe differs from 280 by rounding only
    for W, c in ((Ww, K * D), (We, kap * dTdr), (Wew, el * K * D)):
        assert W[0, 1] == pytest.approx(-c / dz ** 2 - 1, rel=rt)
        assert np.allclose(W[1:-1, 1], -2 * c / dz ** 2 - 1, rtol=rt)
        assert W[-1, 1] == pytest.approx(-c / dz ** 2 - 1, rel=rt)


def test_hydrostatic_kat(oracle_mod):
    """soiltest.jl:15-90: a hydrostatic profile has ψ + z = const and zero
    tendency (zero-flux BCs)."""
    nu, Ks, Ss, n, a, thr = 0.495, 0.0443 / 3600 / 100, 1e-3, 2.0, 2.6, 0.0
    m = 1 - 1 / n
    zf = uniform_column(-10.0, 0.0, 50)
    o = single_column(oracle_mod, 0, 0, zf, [nu, thr, a, n, Ks, Ss])
    zc = 0.5 * (zf[1:] + zf[:-1])
    S = (1 + (a * (zc + 10.0)) ** n) ** (-m)
    Y = o.new_state()
    Y["vartheta_l"][:] = S * (nu - thr) + thr
    aux = o.new_aux()
    z = np.zeros(1)
    o.update_cache(Y, aux, z, None, z)
    dY = o.imp_tendency(Y, aux)
    assert np.mean(dY["vartheta_l"]) < 2.220446049250313e-16
    assert np.abs(dY["vartheta_l"]).max() < 1e-18
    assert abs(np.mean(aux["psi"] + zc) + 10.0) < 2 * 2.220446049250313e-16 * 10


def _stretched_batch(name, ncol=6):
    from clmsoil import inputs
    return inputs.make_batch(name, ncol=ncol)


def test_rre_jacobian_matches_finite_differences(oracle_mod):
    """W = dtγ ∂T/∂ϑ - I with K frozen (modified Picard) on stretched,
    uneven columns: central differences of the oracle's own tendency."""
    b = _stretched_batch("rre_131072_uneven", 4)
    o = oracle_mod.OracleModel.from_batch(b)
    Y = b.copy_state()
    aux = o.new_aux()
    o.update_cache(Y, aux, b.top_water[0].copy())
    dt = 1800.0
    W = o.jacobian(Y, aux, dt)[0].reshape(-1, 3)
    for c in range(b.ncol):
        lo, hi = b.col_ptr[c], b.col_ptr[c + 1]
        N = hi - lo
        J = np.zeros((N, N))
        for j in range(N):
            cols = []
            for sgn in (1, -1):
                Yp = {k: v.copy() for k, v in Y.items()}
                h = 1e-6 * max(abs(Y["vartheta_l"][lo + j]), 1e-3)
                Yp["vartheta_l"][lo + j] += sgn * h
                a2 = o.new_aux()
                o.update_cache(Yp, a2, b.top_water[0].copy())
                a2["K"][:] = aux["K"]          # K frozen
                a2["bot_water"][:] = aux["bot_water"]
                cols.append(o.imp_tendency(Yp, a2)["vartheta_l"][lo:hi])
            J[:, j] = (cols[0] - cols[1]) / (2 * h)
        Wfd = dt * J - np.eye(N)
        Wd = np.diag(W[lo:hi, 1]) + np.diag(W[lo + 1:hi, 0], -1) + np.diag(W[lo:hi - 1, 2], 1)
        assert np.abs(Wd - Wfd).max() <= 1e-6 * np.abs(Wd).max()


def _fd_block(o, Y, aux, key_in, key_out, lo, hi, refresh, dt):
    N = hi - lo
    J = np.zeros((N, N))
    for j in range(N):
        h = 1e-6 * max(abs(Y[key_in][lo + j]), 1e-3)
        cols = []
        for sgn in (1, -1):
            Yp = {k: v.copy() for k, v in Y.items()}
            Yp[key_in][lo + j] += sgn * h
            cols.append(o.imp_tendency(Yp, refresh(Yp))[key_out][lo:hi])
        J[:, j] = (cols[0] - cols[1]) / (2 * h)
    return dt * J - np.eye(N)


def _dense(Wb, lo, hi):
    return np.diag(Wb[lo:hi, 1]) + np.diag(Wb[lo + 1:hi, 0], -1) + np.diag(Wb[lo:hi - 1, 2], 1)


def test_eh_jacobian_matches_finite_differences(oracle_mod):
    """EH blocks against central differences of the oracle's own tendency,
    with the reference's lagging: K, κ, θ_l frozen at u_n; ∂/∂ϑ through ψ
    (ν_eff = ν - θ_i) with ρe_l(T)K frozen, which makes the ρe-ϑ block
    dtγ J - I as the reference assembles it (energy_hydrology.jl:545-556);
    the ρe-ρe block on columns with K_sat -> 0 (pure conduction, where the
    reference's block is exact)."""
    b = _stretched_batch("eh_65536x20", 3)
    o = oracle_mod.OracleModel.from_batch(b)
    Y = b.copy_state()
    aux = o.new_aux()
    o.update_cache(Y, aux, b.top_water[0].copy(), b.top_heat[0].copy())
    dt = 900.0
    Ww, We, Wew = (w.reshape(-1, 3) for w in o.jacobian(Y, aux, dt))

    def psi_only(Yp):
        a2 = {k: v.copy() for k, v in aux.items()}
        o.update_implicit_cache(Yp, a2)
        a2["T"][:] = aux["T"]
        return a2

    for c in range(b.ncol):
        lo, hi = b.col_ptr[c], b.col_ptr[c + 1]
        for Wb, ki, ko in ((Ww, "vartheta_l", "vartheta_l"), (Wew, "vartheta_l", "rho_e_int")):
            Wfd = _fd_block(o, Y, aux, ki, ko, lo, hi, psi_only, dt)
            Wd = _dense(Wb, lo, hi)
            assert np.abs(Wd - Wfd).max() <= 1e-5 * np.abs(Wd).max(), (ki, ko, c)
    # conduction block
    P = b.params.copy()
    P[4] = 1e-30
    o2 = oracle_mod.OracleModel(b.model, b.bc_bottom, b.col_ptr, b.z_face, P)
    aux2 = o2.new_aux()
    o2.update_cache(Y, aux2, b.top_water[0].copy(), b.top_heat[0].copy())
    We2 = o2.jacobian(Y, aux2, dt)[1].reshape(-1, 3)

    def refresh_T(Yp):
        a2 = {k: v.copy() for k, v in aux2.items()}
        o2.update_implicit_cache(Yp, a2)
        return a2

    for c in range(b.ncol):
        lo, hi = b.col_ptr[c], b.col_ptr[c + 1]
        Wfd = _fd_block(o2, Y, aux2, "rho_e_int", "rho_e_int", lo, hi, refresh_T, dt)
        Wd = _dense(We2, lo, hi)
        assert np.abs(Wd - Wfd).max() <= 1e-6 * np.abs(Wd).max(), c


def test_thomas_matches_dense_solve(oracle_mod):
    b = _stretched_batch("eh_262144_uneven", 5)
    o = oracle_mod.OracleModel.from_batch(b)
    Y = b.copy_state()
    aux = o.new_aux()
    o.update_cache(Y, aux, b.top_water[0].copy(), b.top_heat[0].copy())
    W = o.jacobian(Y, aux, 1800.0)
    rng = np.random.default_rng(0)
    R = o.new_state()
    for k in R:
        R[k][:] = rng.standard_normal(R[k].size)
    dx = o.solve(W, R)
    Ww, We, Wew = (w.reshape(-1, 3) for w in W)

    def dense(Wb, lo, hi):
        return np.diag(Wb[lo:hi, 1]) + np.diag(Wb[lo + 1:hi, 0], -1) + np.diag(Wb[lo:hi - 1, 2], 1)

    for c in range(b.ncol):
        lo, hi = b.col_ptr[c], b.col_ptr[c + 1]
        A = np.block([[dense(Ww, lo, hi), np.zeros((hi - lo, hi - lo))],
                      [dense(Wew, lo, hi), dense(We, lo, hi)]])
        rhs = np.concatenate([R["vartheta_l"][lo:hi], R["rho_e_int"][lo:hi]])
        x = np.linalg.solve(A, rhs)
        xo = np.concatenate([dx["vartheta_l"][lo:hi], dx["rho_e_int"][lo:hi]])
        # backward error of the Thomas/block solve, and forward agreement
        # with LU to within the block's condition number
        bwd = np.abs(A @ xo - rhs).max() / (np.abs(A).max() * np.abs(xo).max())
        assert bwd <= 1e-15
        cond = np.linalg.cond(A)
        assert np.abs(xo - x).max() <= 1e-15 * cond * np.abs(x).max()
    assert np.array_equal(dx["theta_i"], -R["theta_i"])
    assert np.array_equal(dx["intF_water"], -R["intF_water"])


def test_tendency_is_conservative(oracle_mod):
    """Σ Δz T_ϑ = -(F_top - F_bot) and Σ Δz T_ρe = -(G_top - G_bot) exactly
    (flux form), which is the discrete statement of conservation."""
    for name in ("rre_131072_uneven", "eh_262144_uneven"):
        b = _stretched_batch(name, 8)
        o = oracle_mod.OracleModel.from_batch(b)
        Y = b.copy_state()
        aux = o.new_aux()
        o.update_cache(Y, aux, b.top_water[0].copy(),
                       None if b.top_heat is None else b.top_heat[0].copy())
        d = o.imp_tendency(Y, aux)
        dz = np.concatenate([np.diff(b.z_face[b.col_ptr[c] + c:b.col_ptr[c + 1] + c + 1])
                             for c in range(b.ncol)])
        sw = np.add.reduceat(dz * d["vartheta_l"], b.col_ptr[:-1])
        assert np.allclose(sw, -(aux["top_water"] - aux["bot_water"]), rtol=1e-9, atol=1e-18)
        if o.eh:
            se = np.add.reduceat(dz * d["rho_e_int"], b.col_ptr[:-1])
            assert np.allclose(se, -(aux["top_heat"] - aux["bot_heat"]), rtol=1e-8, atol=1e-9)


def test_every_newton_iterate_conserves_water(oracle_mod):
    """With an exact linear solve, Σ Δz x_k = Σ Δz U0 - Δt (F_top - F_bot)
    after every iteration, converged or not (4 fixed iterations here).
    Checked on column-steps that stay above θ_r: below it the clipped
    saturation makes dψ/dϑ astronomically large, W is scaled beyond FP64
    and W Δ = R holds only to rounding (the reference shares this)."""
    b = _stretched_batch("rre_131072_uneven", 32)
    w = b.workload
    o = oracle_mod.OracleModel.from_batch(b)
    Y = b.copy_state()
    dz = np.concatenate([np.diff(b.z_face[b.col_ptr[c] + c:b.col_ptr[c + 1] + c + 1])
                         for c in range(b.ncol)])
    thr = np.repeat(b.params[1], b.nlayers)
    nchk = 0
    for s in range(12):
        w0 = np.add.reduceat(dz * Y["vartheta_l"], b.col_ptr[:-1])
        f0 = Y["intF_water"].copy()
        valid0 = np.minimum.reduceat(Y["vartheta_l"] - thr, b.col_ptr[:-1]) > 1e-3
        o.step(Y, w.dt, 4, 0.0, b.top_water[b.forcing_slot(s)].copy())
        w1 = np.add.reduceat(dz * Y["vartheta_l"], b.col_ptr[:-1])
        ok = valid0 & (np.minimum.reduceat(Y["vartheta_l"] - thr, b.col_ptr[:-1]) > 1e-3)
        nchk += ok.sum()
        assert np.allclose((w1 - w0)[ok], (Y["intF_water"] - f0)[ok], rtol=1e-9, atol=1e-12)
    assert nchk > 0.8 * 12 * b.ncol


def test_phase_change_conserves_water_mass(oracle_mod):
    """soiltest.jl:649-737: the explicit phase-change tendencies satisfy
    ρ_l dϑ_l + ρ_i dθ_i = 0 and leave ρe_int unchanged."""
    K = oracle_mod.constants()
    b = _stretched_batch("eh_65536x20", 16)
    o = oracle_mod.OracleModel.from_batch(b)
    Y = b.copy_state()
    aux = o.new_aux()
    o.update_cache(Y, aux, b.top_water[0].copy(), b.top_heat[0].copy())
    d = o.exp_tendency(Y, aux)
    assert np.abs(d["theta_i"]).max() > 0
    assert np.allclose(K["RHO_L"] * d["vartheta_l"] + K["RHO_I"] * d["theta_i"], 0.0,
                       atol=1e-12 * np.abs(d["theta_i"]).max() * K["RHO_L"])
    assert np.all(d["rho_e_int"] == 0.0)


@pytest.mark.parametrize("name", ["rre_4096x50", "eh_65536x20"])
def test_converged_step_is_root_of_implicit_euler(oracle_mod, name):
    """Brute force on tiny columns: a general dense Newton (scipy fsolve on
    R(x) = U0 + Δt T_imp(x) - x with the reference's lagging) and the
    oracle's modified-Picard step converged to 1e-12 reach the same x."""
    from scipy.optimize import fsolve
    from clmsoil import inputs
    b0 = inputs.make_batch(name, ncol=8)
    checked = 0
    for c in range(8):
        sub = b0.column_subset([c])
        o = oracle_mod.OracleModel.from_batch(sub)
        Y = sub.copy_state()
        tw = sub.top_water[0].copy()
        th = None if sub.top_heat is None else sub.top_heat[0].copy()
        aux = o.new_aux()
        o.update_cache(Y, aux, tw, th)
        dt = sub.workload.dt
        U0 = {k: v.copy() for k, v in Y.items()}
        if o.eh:
            e = o.exp_tendency(Y, aux)
            U0 = {k: Y[k] + dt * e[k] for k in Y}
        Ys = {k: v.copy() for k, v in Y.items()}
        it, nc = o.step(Ys, dt, 200, 1e-12, tw, th)
        if nc:
            continue
        N = sub.ncell

        def resid(z):
            x = {k: v.copy() for k, v in U0.items()}
            x["vartheta_l"][:] = z[:N]
            if o.eh:
                x["rho_e_int"][:] = z[N:] * 1e6
            a2 = {k: v.copy() for k, v in aux.items()}
            o.update_implicit_cache(x, a2)
            if not o.eh:
                a2["bot_water"][:] = aux["bot_water"]  # lagged free-drainage flux
            T = o.imp_tendency(x, a2)
            r = U0["vartheta_l"] + dt * T["vartheta_l"] - x["vartheta_l"]
            if o.eh:
                r = np.concatenate([r, (U0["rho_e_int"] + dt * T["rho_e_int"] - x["rho_e_int"]) / 1e6])
            return r

        z0 = U0["vartheta_l"].copy() if not o.eh else \
            np.concatenate([U0["vartheta_l"], U0["rho_e_int"] / 1e6])
        z, info, ier, msg = fsolve(resid, z0, full_output=True, xtol=1e-14)
        if ier != 1:
            continue
        assert np.allclose(Ys["vartheta_l"], z[:N], rtol=1e-9, atol=1e-11), (name, c)
        if o.eh:
            assert np.allclose(Ys["rho_e_int"], z[N:] * 1e6, rtol=1e-9, atol=1e-3), (name, c)
        checked += 1
    assert checked >= 3


@pytest.mark.parametrize("name", ["rre_131072_uneven", "eh_262144_uneven"])
def test_warm_start_reaches_the_same_root(oracle_mod, name):
    """Newton from x0 = U0 + guess: the exact increment converges at once; a
    perturbed guess ends at the plain step's root to the Newton tolerance
    (Newton still decides convergence).  For RichardsModel a zero guess is the plain
    step bit for bit; for EnergyHydrology it is not, because the warm start
    refreshes T and ψ at x0 while the plain first iteration uses those of
    u_n -- it still reaches the same root."""
    from clmsoil import inputs
    b = inputs.make_batch(name, ncol=256)
    w = b.workload
    o = oracle_mod.OracleModel.from_batch(b)
    fw = b.top_water[0].copy()
    fh = None if b.top_heat is None else b.top_heat[0].copy()
    eh = b.model == 1
    U = b.copy_state()
    o.step(U, w.dt, 0, w.newton_tol, fw, fh)  # zero iterations: the predictor U0
    X = b.copy_state()
    it_x, _ = o.step(X, w.dt, 200, 1e-13, fw, fh)
    P = b.copy_state()
    it_p, _ = o.step(P, w.dt, w.newton_max_iters, w.newton_tol, fw, fh)
    Z = b.copy_state()
    zero = np.zeros(b.ncell)
    it_z, _ = o.step(Z, w.dt, w.newton_max_iters, w.newton_tol, fw, fh, guess_v=zero,
                     guess_e=zero if eh else None)
    if not eh:
        assert np.array_equal(it_z, it_p)
        for k in P:
            assert np.array_equal(np.nan_to_num(Z[k]), np.nan_to_num(P[k])), k
    good = (it_x < 200) & (it_p < w.newton_max_iters)
    good &= np.isfinite(np.add.reduceat(X["vartheta_l"], b.col_ptr[:-1]))
    assert good.mean() > 0.5
    cm = np.repeat(good, b.nlayers)
    gv = np.where(cm, X["vartheta_l"] - U["vartheta_l"], 0.0)
    ge = np.where(cm, X["rho_e_int"] - U["rho_e_int"], 0.0) if eh else None
    E = b.copy_state()
    it_e, _ = o.step(E, w.dt, w.newton_max_iters, w.newton_tol, fw, fh, guess_v=gv, guess_e=ge)
    assert (it_e[good] <= 2).all()
    assert np.abs(E["vartheta_l"] - X["vartheta_l"])[cm].max() <= 1e-9
    rng = np.random.default_rng(3)
    Q = b.copy_state()
    it_q, _ = o.step(Q, w.dt, w.newton_max_iters, w.newton_tol, fw, fh,
                     guess_v=gv * (1 + 0.1 * rng.standard_normal(b.ncell)),
                     guess_e=None if ge is None else ge * (1 + 0.1 * rng.standard_normal(b.ncell)))
    both = good & (it_q < w.newton_max_iters)
    cb = np.repeat(both, b.nlayers)
    assert both.mean() > 0.5
    # two Newton solutions to tolerance 1e-8 from different starts agree to ~ the tolerance
    assert np.abs(Q["vartheta_l"] - X["vartheta_l"])[cb].max() <= 1e-7
    assert np.abs(P["vartheta_l"] - X["vartheta_l"])[cb].max() <= 1e-7
    cz = np.repeat(good & (it_z < w.newton_max_iters), b.nlayers)
    assert np.abs(Z["vartheta_l"] - X["vartheta_l"])[cz].max() <= 1e-7
