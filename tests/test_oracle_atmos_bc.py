"""Oracle AtmosDrivenFluxBC for EnergyHydrology in soil-only form
(boundary_conditions.jl:646-1003, energy_hydrology.jl:909-945), with the
turbulent fluxes and net radiation supplied per column, and NoRunoff or
TOPMODELRunoff (Runoff/Runoff.jl:234-359).

Pinned by the flux assembly on designed columns (top water = infiltration +
vapor_flux_liq, top heat = R_n + lhf + shf + infiltration ρe_l(T_air)); the
EH TOPMODEL closed forms where the surface pass counts ice and the
subsurface pass does not (a layer saturated by ice only changes the
infiltration, not R_ss), the EH infiltration capacity through the already
pinned impedance and viscosity factors, and R_ess as the liquid energy of
the saturated layers; the sublimation source on the top layer only (and
the reference's use of the absolute height z, which reaches further down
when the surface is above z = 0); the implicit tendency's difference from a
flux BC (exactly the sinks); and the water and energy balance of a step.
"""
import math

import numpy as np
import pytest

from oracle.oracle import (BC_ATMOS, BC_ATMOS_TOPMODEL, BC_FLUX, BC_FREE_DRAINAGE, OracleModel,
                           constants, lib)

CS = constants()
RHO_L, RHO_I, CP_L, T_0, LH_F0 = CS["RHO_L"], CS["RHO_I"], CS["CP_L"], CS["T_0"], CS["LH_F0"]
F_OVER, R_SB = 3.28, 1.484e-7
NU, THR, KSAT, RHOC_DS = 0.45, 0.05, 1e-6, 2e6


def rho_e_liq(T):
    return RHO_L * CP_L * (T - T_0)


def column(zf, bc_top=BC_ATMOS, f_max=0.5):
    rows = [NU, THR, 2.0, 1.6, KSAT, 1e-3, 0.0, 0.5, 0.0, 0.25, 2.3, 1.5, RHOC_DS]
    if bc_top == BC_ATMOS_TOPMODEL:
        rows += [f_max, F_OVER, R_SB]
    N = zf.size - 1
    return OracleModel(1, BC_FREE_DRAINAGE, np.array([0, N], dtype=np.int32), zf,
                       np.asarray(rows, dtype=np.float64).reshape(-1, 1), bc_top=bc_top)


def state(o, vl, ti, T):
    """EH state with temperature T (K) per layer: ρe = ρc_s (T - T_0) - θ_i ρ_i L_f."""
    Y = o.new_state()
    Y["vartheta_l"][:] = vl
    Y["theta_i"][:] = ti
    thl = np.minimum(vl, NU - ti)
    rcs = RHOC_DS + thl * RHO_L * CP_L + ti * RHO_I * CS["CP_I"]
    Y["rho_e_int"][:] = rcs * (T - T_0) - ti * RHO_I * LH_F0
    return Y


def sfc(P=-2e-5, T_air=285.0, lhf=-80.0, shf=-20.0, vliq=3e-8, vice=0.0, R_n=-300.0):
    return {k: np.array([v]) for k, v in dict(P_liq=P, T_air=T_air, lhf=lhf, shf=shf,
                                             vapor_flux_liq=vliq, vapor_flux_ice=vice,
                                             R_n=R_n).items()}


def cache(o, Y, s):
    a = o.new_aux()
    o.update_cache(Y, a, None, None, None, None, sfc=s)
    return a


def test_flux_assembly_no_runoff():
    o = column(np.linspace(-2.0, 0.0, 21))
    Y = state(o, np.full(20, 0.3), np.zeros(20), np.full(20, 280.0))
    s = sfc()
    a = cache(o, Y, s)
    assert a["top_water"][0] == pytest.approx(-2e-5 + 3e-8, rel=1e-15)
    assert a["top_heat"][0] == pytest.approx(-300.0 - 80.0 - 20.0 + -2e-5 * rho_e_liq(285.0),
                                             rel=1e-15)
    assert a["vapor_ice"][0] == 0.0 and (a["sink"] == 0).all()


def test_topmodel_eh_ice_counts_for_the_surface_only():
    """2 m, 20 layers of 0.1 m.  Bottom four layers hold liquid above ν
    (weight 1.1 in both passes); the top layer holds ϑ_l = 0.3, θ_i = 0.2
    (total weight 1.125 in the surface pass, 0 in the subsurface pass)."""
    L = lib()
    o = column(np.linspace(-2.0, 0.0, 21), BC_ATMOS_TOPMODEL)
    vl = np.full(20, 0.3)
    vl[:4] = THR + 1.1 * (NU - THR)
    ti = np.zeros(20)
    ti[-1] = 0.2
    T = np.linspace(276.0, 272.0, 20)
    Y = state(o, vl, ti, T)
    a = cache(o, Y, sfc(P=-1e-3))   # P below the capacity: infiltration at i_c
    hs = 4 * 0.1 * 1.1 + 0.1 * 1.125
    h = 4 * 0.1 * 1.1
    thl_top = min(0.3, NU - 0.2)
    ic = -KSAT * L.orc_impedance_factor(0.2 / (thl_top + 0.2 - THR), CS["OMEGA"]) * \
        L.orc_viscosity_factor(a["T"][-1], CS["GAMMA_VISC"], CS["GAMMA_T_REF"])
    fsat = 0.5 * math.exp(-F_OVER / 2 * (2.0 - hs))
    infil = (1 - fsat) * ic
    assert a["top_water"][0] == pytest.approx(infil + 3e-8, rel=1e-12)
    rss = R_SB * math.exp(-F_OVER * (2.0 - h))
    assert a["r_ss"][0] == pytest.approx(rss, rel=1e-12)
    he = sum(1.1 * rho_e_liq(a["T"][i]) * 0.1 for i in range(4))
    ress = he * rss / h
    assert a["r_ess"][0] == pytest.approx(ress, rel=1e-12)
    np.testing.assert_allclose(a["sink"][:4], rss / h * 1.1, rtol=1e-12)
    np.testing.assert_allclose(a["sink_e"][:4], ress / h * 1.1, rtol=1e-12)
    assert (a["sink"][4:] == 0).all() and (a["sink_e"][4:] == 0).all()
    assert a["top_heat"][0] == pytest.approx(-400.0 + infil * rho_e_liq(285.0), rel=1e-12)


def test_sublimation_top_layer_and_absolute_height():
    zf = -2.0 * (1.05 ** np.arange(20, -1, -1) - 1) / (1.05 ** 20 - 1)
    for shift, nlayers in ((0.0, 1), (0.5, None)):
        o = column(zf + shift)
        Y = state(o, np.full(20, 0.3), np.full(20, 0.05), np.full(20, 268.0))
        a = cache(o, Y, sfc(vice=2e-8))
        d = o.exp_tendency(Y, a)
        o0 = column(zf + shift, BC_FLUX)
        a0 = o0.new_aux()
        o0.update_cache(Y, a0, np.zeros(1), np.zeros(1))
        d0 = o0.exp_tendency(Y, a0)
        extra = d["theta_i"] - d0["theta_i"]
        dzt = 0.5 * (zf[-1] - zf[-2])
        zc = 0.5 * (zf[1:] + zf[:-1]) + shift
        expect = np.where(zc + 2 * dzt > np.finfo(float).eps,
                          -2e-8 * RHO_L / RHO_I / (2 * dzt), 0.0)
        np.testing.assert_allclose(extra, expect, rtol=1e-12, atol=1e-30)
        if nlayers is not None:
            assert (extra != 0).sum() == nlayers
        else:  # surface at z = 0.5: H(z + 2Δz_top) holds for every layer above -2Δz_top
            assert (extra != 0).sum() > 1
        assert d["intF_water"][0] - d0["intF_water"][0] == pytest.approx(-2e-8, rel=1e-14)


def test_imp_tendency_is_flux_tendency_minus_sinks():
    zf = np.linspace(-2.0, 0.0, 21)
    o = column(zf, BC_ATMOS_TOPMODEL)
    vl = np.linspace(0.47, 0.25, 20)
    Y = state(o, vl, np.zeros(20), np.linspace(285.0, 278.0, 20))
    a = cache(o, Y, sfc())
    assert (a["sink"] > 0).any() and (a["sink"] == 0).any()
    d = o.imp_tendency(Y, a)
    of = column(zf, BC_FLUX)
    af = of.new_aux()
    of.update_cache(Y, af, a["top_water"].copy(), a["top_heat"].copy())
    df = of.imp_tendency(Y, af)
    np.testing.assert_allclose(d["vartheta_l"], df["vartheta_l"] - a["sink"], rtol=1e-13,
                               atol=1e-22)
    np.testing.assert_allclose(d["rho_e_int"], df["rho_e_int"] - a["sink_e"], rtol=1e-13,
                               atol=1e-12)
    assert d["intF_water"][0] == pytest.approx(df["intF_water"][0] - a["r_ss"][0], rel=1e-14)
    assert d["intF_energy"][0] == pytest.approx(df["intF_energy"][0] - a["r_ess"][0], rel=1e-14)


@pytest.mark.parametrize("bc", [BC_ATMOS, BC_ATMOS_TOPMODEL])
def test_step_water_and_energy_balance(bc):
    """Σ Δz (Δϑ_l + ρ_i/ρ_l Δθ_i) = Δ∫F_w and Σ Δz Δρe = Δ∫F_e over a step
    with sublimation, runoff and phase change (which conserves both)."""
    zf = np.linspace(-2.0, 0.0, 31)
    o = column(zf, bc, f_max=0.3)
    zc = 0.5 * (zf[1:] + zf[:-1])
    vl = np.where(zc < -1.4, NU + 1e-3 * (-1.4 - zc), 0.28)
    ti = np.where(zc > -0.3, 0.05, 0.0)
    Y = state(o, vl, ti, np.where(zc > -0.3, 271.0, 276.0))
    Y0 = {k: v.copy() for k, v in Y.items()}
    s = sfc(P=-1e-6, vice=1e-8)
    it, _ = o.step(Y, 900.0, 50, 1e-12, None, None, None, None, sfc=s)
    assert it[0] < 50
    dz = np.diff(zf)
    dw = ((Y["vartheta_l"] - Y0["vartheta_l"]) +
          RHO_I / RHO_L * (Y["theta_i"] - Y0["theta_i"])) @ dz
    de = (Y["rho_e_int"] - Y0["rho_e_int"]) @ dz
    assert dw == pytest.approx(Y["intF_water"][0] - Y0["intF_water"][0], rel=1e-7, abs=1e-12)
    assert de == pytest.approx(Y["intF_energy"][0] - Y0["intF_energy"][0], rel=1e-7, abs=1e-3)
    assert abs(Y["intF_water"][0]) > 0 and abs(Y["intF_energy"][0]) > 0
