"""Oracle MoistureStateBC (RichardsModel): the reference's Jacobian KAT with
a top moisture state (test/shared_utilities/implicit_timestepping/
richards_model.jl:16-145), a hydrostatic column held by state BCs at both
ends (zero boundary flux pins sign and gravity), finite differences of the
top row, and the converged step as the root of the implicit residual with
the boundary fluxes refreshed at the iterate (rre.jl:460-468)."""
import numpy as np
import pytest

from oracle.oracle import BC_FREE_DRAINAGE, BC_MOISTURE_STATE, OracleModel


def column(zf, params, bc_top=BC_MOISTURE_STATE, bc_bottom=BC_FREE_DRAINAGE):
    N = zf.size - 1
    return OracleModel(0, bc_bottom, np.array([0, N], dtype=np.int32), zf,
                       np.asarray(params, dtype=np.float64).reshape(-1, 1), bc_top=bc_top)


P = dict(nu=0.495, thr=0.124, a=2.6, n=1.43, Ksat=0.0443 / 3600 / 100, Ss=1e-3)
PARAMS = [P["nu"], P["thr"], P["a"], P["n"], P["Ksat"], P["Ss"]]


def test_jacobian_kat_moisture_state_top(oracle_mod):
    """richards_model.jl:16-145: 150 uniform layers on [-1.5, 0], ϑ = 0.24,
    top ϑ_bc = ν - 1e-3, free drainage, dtγ = 1, after cache_imp!."""
    o = column(np.linspace(-1.5, 0.0, 151), PARAMS)
    Y = o.new_state()
    Y["vartheta_l"][:] = 0.24
    aux = o.new_aux()
    o.update_implicit_cache(Y, aux, np.array([P["nu"] - 1e-3]), np.array([0.0]))
    W = o.jacobian(Y, aux, 1.0)[0].reshape(-1, 3)
    L = oracle_mod.lib()
    S = L.orc_effective_saturation(P["nu"], 0.24, P["thr"])
    K = L.orc_hydraulic_conductivity(P["n"], P["Ksat"], S)
    D = L.orc_dpsidtheta(P["a"], P["n"], 0.24, P["nu"], P["thr"], P["Ss"])
    dz = 0.01
    assert W[0, 0] == 0.0 and W[-1, 2] == 0.0
    assert np.allclose(W[1:, 0], W[:-1, 2], rtol=1e-12)
    assert np.allclose(W[1:, 0], K / dz ** 2 * D, rtol=1e-10)
    assert np.isclose(W[0, 1], -K / dz ** 2 * D - 1, rtol=1e-10)
    assert np.allclose(W[1:-1, 1], -2 * K / dz ** 2 * D - 1, rtol=1e-10)
    assert np.isclose(W[-1, 1], -K / dz ** 2 * D - K / (dz * dz / 2) * D - 1, rtol=1e-10)


def _hydrostatic(zf, psi_top):
    """ϑ with ψ + z constant, and the boundary states that continue it."""
    L = oracle_mod_lib()
    zc = 0.5 * (zf[1:] + zf[:-1])
    h = psi_top + zc[-1]

    def theta(psi):
        import ctypes as C
        e = C.c_int(0)
        S = L.orc_inverse_matric_potential(P["a"], P["n"], psi, C.byref(e))
        return P["thr"] + (P["nu"] - P["thr"]) * S

    v = np.array([theta(h - z) for z in zc])
    return v, theta(h - zf[-1]), theta(h - zf[0])


def oracle_mod_lib():
    from oracle.oracle import lib
    return lib()


@pytest.mark.parametrize("bc_bottom", [BC_MOISTURE_STATE])
def test_hydrostatic_column_held_by_state_bcs(oracle_mod, bc_bottom):
    zf = np.sort(np.concatenate([[-2.0], -np.geomspace(1.9, 0.02, 29), [0.0]]))
    o = column(zf, PARAMS, BC_MOISTURE_STATE, bc_bottom)
    v, top, bot = _hydrostatic(zf, -3.0)  # unsaturated throughout
    Y = o.new_state()
    Y["vartheta_l"][:] = v
    aux = o.new_aux()
    o.update_cache(Y, aux, np.array([top]), None, np.array([bot]))
    scale = P["Ksat"]
    assert abs(aux["top_water"][0]) <= 1e-9 * scale
    assert abs(aux["bot_water"][0]) <= 1e-9 * scale
    d = o.imp_tendency(Y, aux)
    assert np.abs(d["vartheta_l"]).max() <= 1e-9 * scale
    # a wetter top state drives water down (negative flux), a drier one up
    o.update_cache(Y, aux, np.array([top + 0.05]), None, np.array([bot]))
    assert aux["top_water"][0] < 0
    o.update_cache(Y, aux, np.array([top - 0.05]), None, np.array([bot]))
    assert aux["top_water"][0] > 0
    # a wetter bottom state pushes water up
    o.update_cache(Y, aux, np.array([top]), None, np.array([min(bot + 0.05, P["nu"])]))
    assert aux["bot_water"][0] > 0


def test_top_row_matches_finite_differences(oracle_mod):
    """d(T_N)/dϑ_N with K frozen, ψ and the top boundary flux at the iterate."""
    zf = np.concatenate([np.linspace(-1.0, -0.2, 9), [-0.1, -0.04, 0.0]])
    o = column(zf, PARAMS)
    N = zf.size - 1
    Y = o.new_state()
    Y["vartheta_l"][:] = np.linspace(0.30, 0.40, N)
    tb = np.array([0.45])
    aux = o.new_aux()
    o.update_cache(Y, aux, tb, None, np.array([0.0]))
    dt = 1800.0
    W = o.jacobian(Y, aux, dt)[0].reshape(-1, 3)
    L = oracle_mod.lib()
    zc = 0.5 * (zf[1:] + zf[:-1])
    dztop = zf[-1] - zc[-1]
    psib = L.orc_pressure_head(P["a"], P["n"], P["thr"], tb[0], P["nu"], P["Ss"])

    def tend(x):
        Yp = {k: v.copy() for k, v in Y.items()}
        Yp["vartheta_l"][:] = x
        a2 = o.new_aux()
        o.update_cache(Yp, a2, tb, None, np.array([0.0]))
        a2["K"][:] = aux["K"]
        a2["bot_water"][:] = aux["bot_water"]
        a2["top_water"][:] = -aux["K"][-1] * ((psib + dztop) - a2["psi"][-1]) / dztop
        return o.imp_tendency(Yp, a2)["vartheta_l"]

    x = Y["vartheta_l"].copy()
    J = np.zeros((N, N))
    for j in range(N):
        h = 1e-6 * x[j]
        e = np.zeros(N)
        e[j] = h
        J[:, j] = (tend(x + e) - tend(x - e)) / (2 * h)
    Wfd = dt * J - np.eye(N)
    Wd = np.diag(W[:, 1]) + np.diag(W[1:, 0], -1) + np.diag(W[:-1, 2], 1)
    assert np.abs(Wd - Wfd).max() <= 1e-6 * np.abs(Wd).max()
    # the boundary term is what distinguishes the top diagonal
    assert abs(W[-1, 1] - Wfd[-1, -1]) <= 1e-6 * abs(W[-1, 1])


def test_converged_step_is_root_with_implicit_bc(oracle_mod):
    """Newton with the boundary fluxes refreshed at every iterate converges
    to the root of R(x) = u + Δt T(x; F_bc(x)) - x found by scipy."""
    from scipy.optimize import fsolve
    zf = np.concatenate([np.linspace(-1.5, -0.3, 7), [-0.15, -0.05, 0.0]])
    o = column(zf, PARAMS)
    N = zf.size - 1
    u = {"vartheta_l": np.linspace(0.28, 0.33, N), "intF_water": np.zeros(1)}
    tb, bb = np.array([0.47]), np.array([0.0])
    dt = 900.0
    Ys = {k: v.copy() for k, v in u.items()}
    it, nc = o.step(Ys, dt, 200, 1e-12, tb, None, bb)
    assert nc == 0

    def resid(z):
        x = {"vartheta_l": z.copy(), "intF_water": np.zeros(1)}
        a = o.new_aux()
        o.update_cache(x, a, tb, None, bb)
        T = o.imp_tendency(x, a)
        return u["vartheta_l"] + dt * T["vartheta_l"] - z

    z, info, ier, msg = fsolve(resid, u["vartheta_l"].copy(), full_output=True, xtol=1e-14)
    assert ier == 1
    assert np.allclose(Ys["vartheta_l"], z, rtol=1e-9, atol=1e-12)
    # ∫F closes the water balance with the final boundary fluxes
    a = o.new_aux()
    o.update_cache(Ys, a, tb, None, bb)
    dzc = np.diff(zf)
    dW = (Ys["vartheta_l"] - u["vartheta_l"]) @ dzc
    assert np.isclose(dW, -dt * (a["top_water"][0] - a["bot_water"][0]), rtol=1e-7, atol=1e-12)
