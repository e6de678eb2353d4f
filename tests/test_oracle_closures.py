"""Oracle pointwise closures against the reference's own known-answer tests
(test/standalone/Soil/soil_parameterizations.jl) and against mathematics
that fixes them independently of the code under test."""
import math
import re

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def L(oracle_mod):
    return oracle_mod.lib()


@pytest.fixture(scope="module")
def K(oracle_mod):
    return oracle_mod.constants()


def isapprox(a, b, rtol=1.4901161193847656e-08):  # Julia's default ≈ for Float64
    return abs(a - b) <= rtol * max(abs(a), abs(b))


# ---- soil_parameterizations.jl:225-281 (Richards / van Genuchten) ---------

def test_vg_kat(L):
    th_r, nu, Ss, a, n, Ks = 0.2, 0.4, 1e-2, 3.6, 1.56, 2.9e-7
    m = 1.0 - 1.0 / n
    th = [0.3, 0.4, 0.5]
    S = [L.orc_effective_saturation(nu, t, th_r) for t in th]
    assert np.allclose(S, [0.5, 1.0, 1.5], rtol=1e-15)
    va = -((S[0] ** (-1.0 / m) - 1.0) * a ** (-n)) ** (1.0 / n)
    psi = [L.orc_matric_potential(a, n, x) for x in S[:2]]
    assert isapprox(psi[0], va) and psi[1] == 0.0
    err = __import__("ctypes").c_int(0)
    back = [L.orc_inverse_matric_potential(a, n, p, __import__("ctypes").byref(err)) for p in psi]
    assert np.allclose(back, S[:2], rtol=1e-13) and err.value == 0
    dpsi = [L.orc_dpsidtheta(a, n, t, nu, th_r, Ss) for t in th]
    va_d = 1.0 / (a * m * n) / (nu - th_r) * (S[0] ** (-1 / m) - 1) ** (1 / n - 1) * \
        S[0] ** (-1 / m - 1)
    assert isapprox(dpsi[0], va_d) and dpsi[1] == 1 / Ss and dpsi[2] == 1 / Ss
    p = [L.orc_pressure_head(a, n, th_r, t, nu, Ss) for t in th]
    assert isapprox(p[0], psi[0]) and p[1] == 0.0 and isapprox(p[2], 10.0)
    k = [L.orc_hydraulic_conductivity(n, Ks, x) for x in S]
    va_k = math.sqrt(S[0]) * (1 - (1 - S[0] ** (1 / m)) ** m) ** 2 * Ks
    assert isapprox(k[0], va_k) and k[1] == Ks and k[2] == Ks
    vlf = [L.orc_volumetric_liquid_fraction(t, 0.5, 0.0) for t in (0.25, 0.5, 0.75)]
    assert np.allclose(vlf, [0.25, 0.5, 0.5], rtol=0, atol=0)


def test_inverse_matric_raises_for_positive_psi(L):
    import ctypes
    err = ctypes.c_int(0)
    L.orc_inverse_matric_potential(2.0, 1.5, 0.1, ctypes.byref(err))
    assert err.value == 1  # soil_hydrology_parameterizations.jl:73


# ---- soil_parameterizations.jl:9-188 (heat) -------------------------------

def test_heat_kat(L, K):
    rho_l, rho_i = K["RHO_L"], K["RHO_I"]
    rcp_l, rcp_i = K["CP_L"] * rho_l, K["CP_I"] * rho_i
    T_ref, LH_f0 = K["T_0"], K["LH_F0"]
    assert LH_f0 == K["LH_S0"] - K["LH_V0"]  # Parameters.jl:58
    assert (K["KERSTEN_ALPHA"], K["KERSTEN_BETA"], K["GAMMA_VISC"], K["OMEGA"],
            K["GAMMA_T_REF"]) == (0.24, 18.3, 2.64e-2, 7.0, 288.0)
    nu, nu_om, nu_g, nu_q = 0.2, 0.1, 0.1, 0.1
    k_min, k_om, k_q, k_ice, k_liq = 2.5, 0.25, 8.0, 2.21, 0.57
    rp_q, rp_min, rp_om = 2.66e3, 2.65e3, 1.3e3
    rc_q, rc_om, rc_min = 2.01e6, 2.51e6, 2.01e6
    ks = L.orc_kappa_solid(nu_om, nu_q, k_om, k_q, k_min)
    rp = nu_om * rp_om + nu_q * rp_q + nu_g * rp_min + (1 - nu_om - nu_q - nu_g) * rp_min
    rcss = nu_om * rc_om + nu_q * rc_q + nu_g * rc_min + (1 - nu_om - nu_q - nu_g) * rc_min
    rc_ds = (1 - nu) * rcss
    assert L.orc_temperature_from_rho_e_int(5.4e7, 0.05, 2.1415e6) == \
        T_ref + (5.4e7 + 0.05 * rho_i * LH_f0) / 2.1415e6
    assert isapprox(L.orc_volumetric_heat_capacity(0.25, 0.05, rc_ds),
                    rc_ds + 0.25 * rcp_l + 0.05 * rcp_i)
    assert L.orc_volumetric_internal_energy(0.05, 2.1415e6, 300.0) == \
        2.1415e6 * (300.0 - T_ref) - 0.05 * rho_i * LH_f0
    assert isapprox(L.orc_kappa_sat(0.25, 0.05, 0.57, 2.29),
                    0.57 ** (0.25 / 0.3) * 2.29 ** (0.05 / 0.3))
    assert L.orc_kappa_sat(0.0, 0.0, 0.57, 2.29) == 0.5 * (0.57 + 2.29)
    assert L.orc_relative_saturation(0.25, 0.05, 0.4) == (0.25 + 0.05) / 0.4
    ke0 = L.orc_kersten_number(0.0, 0.75, 0.24, 18.3, nu_om, nu_q, nu_g)
    assert isapprox(ke0, 0.75 ** ((1 + 0.1 - 0.24 * 0.1 - 0.1) / 2) *
                    ((1 + math.exp(-18.3 * 0.75)) ** -3 - ((1 - 0.75) / 2) ** 3) ** (1 - 0.1))
    assert L.orc_kersten_number(0.05, 0.75, 0.24, 18.3, nu_om, nu_q, nu_g) == 0.75 ** 1.1
    assert L.orc_thermal_conductivity(1.5, 0.7287, 0.7187) == 0.7287 * 0.7187 + (1 - 0.7287) * 1.5
    assert L.orc_volumetric_internal_energy_liq(300.0) == rcp_l * (300.0 - T_ref)
    assert isapprox(L.orc_kappa_solid(0.5, 0.25, 2.0, 3.0, 2.0), 2 ** 0.5 * 2 ** 0.25 * 3 ** 0.25)
    assert L.orc_kappa_sat_frozen(0.5, 0.1, 0.4) == 0.5 ** 0.9 * 0.4 ** 0.1
    assert L.orc_kappa_sat_unfrozen(0.5, 0.1, 0.4) == 0.5 ** 0.9 * 0.4 ** 0.1
    rb = (1 - nu) * rp
    assert L.orc_kappa_dry(rp, nu, ks, K["K_THERM"]) == \
        ((0.053 * ks - K["K_THERM"]) * rb + K["K_THERM"] * rp) / (rp - (1 - 0.053) * rb)
    assert isapprox(L.orc_impedance_factor(1.0, 7.0), 1e-7)
    for T in (278.0, 288.0, 298.0):
        assert isapprox(L.orc_viscosity_factor(T, 2.64e-2, 288.0), math.exp(2.64e-2 * (T - 288)))
    assert L.orc_heaviside(0.0) == 0.0 and L.orc_heaviside(1e-15) == 1.0  # x > eps (utils.jl:93-99)
    assert L.orc_heaviside(2.220446049250313e-16) == 0.0


# ---- soil_parameterizations.jl:283-495 (freeze/thaw) ----------------------

def _pcs(L, thl, thi, T, tau, nu, thr, a, n):
    import ctypes
    err = ctypes.c_int(0)
    v = L.orc_phase_change_source(thl, thi, T, tau, nu, thr, a, n, ctypes.byref(err))
    assert err.value == 0
    return v


def test_phase_change_kat(L, K):
    nu, thr, a, n = 0.2, 0.1, 2.0, 1.4
    m = 1 - 1 / n
    rho_l, rho_i, g, Lf, Tf0 = K["RHO_L"], K["RHO_I"], K["GRAV"], K["LH_F0"], K["T_FREEZE"]
    rc_ds = 1.7e6  # any positive dry heat capacity; tau only scales the source
    kd = 0.3

    def expected(thl, thi, T, tau):
        tt = min(rho_i / rho_l * thi + thl, nu)
        S0 = (max(tt, thr + 1.4901161193847656e-08) - thr) / (nu - thr)
        psi0 = -((S0 ** (-1 / m) - 1) * a ** (-n)) ** (1 / n)
        Tf = max(Tf0 * math.exp(g * psi0 / Lf), 1.0)
        psiT = Lf / g * math.log(T / Tf) * (1.0 if Tf - T > 2.220446049250313e-16 else 0.0)
        S = (1 + (a * abs(psi0 + psiT)) ** n) ** (-m)
        return (thl - (S * (nu - thr) + thr)) / tau

    def tau(thl, thi):
        return L.orc_thermal_time(L.orc_volumetric_heat_capacity(thl, thi, rc_ds), 1.0, kd)

    # T = 270: freezing everywhere (3 positive sources)
    v = [_pcs(L, t, 0.0, 270.0, tau(t, 0.0), nu, thr, a, n) for t in (0.11, 0.15, nu)]
    assert all(isapprox(x, expected(t, 0.0, 270.0, tau(t, 0.0))) for x, t in zip(v, (0.11, 0.15, nu)))
    assert sum(x > 0 for x in v) == 3
    # T = 274, no ice: no phase change
    v = [_pcs(L, t, 0.0, 274.0, tau(t, 0.0), nu, thr, a, n) for t in (0.11, 0.15, nu)]
    assert np.allclose(v, 0.0, atol=1e-20)
    # T = 274 with ice: melting (2 negative)
    v = [_pcs(L, 0.11, ti, 274.0, tau(0.11, ti), nu, thr, a, n) for ti in (0.05, 0.08)]
    assert sum(x < 0 for x in v) == 2
    # T = 260 with ice: freezing (2 positive)
    v = [_pcs(L, 0.11, ti, 260.0, tau(0.11, ti), nu, thr, a, n) for ti in (0.05, 0.08)]
    assert sum(x > 0 for x in v) == 2
    assert all(isapprox(x, expected(0.11, ti, 260.0, tau(0.11, ti))) for x, ti in zip(v, (0.05, 0.08)))


# ---- independent pins -------------------------------------------------------

@pytest.mark.parametrize("a,n", [(0.5, 1.2), (2.6, 1.43), (5.0, 2.0), (10.0, 3.0)])
def test_mualem_vg_conductivity_in_psi_form(L, a, n):
    """van Genuchten (1980) eq. 8 gives K(h) in terms of the head; with
    S(ψ) from the inverse closure it must equal K(S)."""
    import ctypes
    m = 1 - 1 / n
    err = ctypes.c_int(0)
    for psi in (-1e-3, -0.05, -0.7, -3.0, -25.0):
        S = L.orc_inverse_matric_potential(a, n, psi, ctypes.byref(err))
        ah = a * abs(psi)
        k_h = (1 - ah ** (n - 1) * (1 + ah ** n) ** (-m)) ** 2 / (1 + ah ** n) ** (m / 2)
        assert L.orc_hydraulic_conductivity(n, 1.0, S) == pytest.approx(k_h, rel=1e-9)
        # round trip of the matric potential
        assert L.orc_matric_potential(a, n, S) == pytest.approx(psi, rel=1e-11)


@pytest.mark.parametrize("a,n,nu,thr", [(0.8, 1.3, 0.45, 0.05), (4.0, 2.5, 0.5, 0.1)])
def test_dpsidtheta_is_derivative_of_pressure_head(L, a, n, nu, thr):
    """dψ/dϑ equals the derivative of pressure_head (central differences),
    in both the unsaturated and the S >= 1 branch."""
    Ss = 1e-3
    for th in (thr + 0.02, 0.5 * (thr + nu), nu - 0.01, nu + 0.01):
        h = 1e-7
        fd = (L.orc_pressure_head(a, n, thr, th + h, nu, Ss) -
              L.orc_pressure_head(a, n, thr, th - h, nu, Ss)) / (2 * h)
        assert L.orc_dpsidtheta(a, n, th, nu, thr, Ss) == pytest.approx(fd, rel=1e-6)


def test_kersten_branches_and_limits(L):
    """Kersten number: 0 for dry soil, 1 at full saturation (both branches),
    and increasing with saturation."""
    for thi, f in ((0.0, lambda S: L.orc_kersten_number(0.0, S, 0.24, 18.3, 0.0, 0.5, 0.0)),
                   (0.1, lambda S: L.orc_kersten_number(0.1, S, 0.24, 18.3, 0.0, 0.5, 0.0))):
        assert f(1.0) == pytest.approx(1.0, abs=1e-7)
        vals = [f(s) for s in np.linspace(0.01, 1.0, 50)]
        assert np.all(np.diff(vals) > 0)
    assert L.orc_kersten_number(0.0, 0.0, 0.24, 18.3, 0.0, 0.5, 0.0) == 0.0


def test_energy_temperature_round_trip(L):
    for thl, thi, rc, T in ((0.3, 0.0, 2e6, 280.0), (0.1, 0.2, 1.5e6, 265.0)):
        rcs = L.orc_volumetric_heat_capacity(thl, thi, rc)
        e = L.orc_volumetric_internal_energy(thi, rcs, T)
        assert L.orc_temperature_from_rho_e_int(e, thi, rcs) == pytest.approx(T, rel=1e-14)


def test_constants_agree_across_oracle_kernels_and_generator(K):
    """One set of physical constants: oracle, device header, host generator."""
    hdr = open(f"{ROOT}/climaland.jl_amd/csrc/clm_physics.h").read()
    dev = {m.group(1): float(m.group(2)) for m in
           re.finditer(r"constexpr double (\w+) = ([0-9.eE+-]+);", hdr)}
    for kd, ko in (("RHO_L", "RHO_L"), ("RHO_I", "RHO_I"), ("CP_L", "CP_L"), ("CP_I", "CP_I"),
                   ("T_0", "T_0"), ("T_FREEZE", "T_FREEZE"), ("LH_V0", "LH_V0"),
                   ("LH_S0", "LH_S0"), ("GRAV", "GRAV"), ("KERSTEN_A", "KERSTEN_ALPHA"),
                   ("KERSTEN_B", "KERSTEN_BETA"), ("OMEGA", "OMEGA"),
                   ("GAMMA_VISC", "GAMMA_VISC"), ("GAMMA_T_REF", "GAMMA_T_REF")):
        assert dev[kd] == K[ko], kd
    from clmsoil import inputs
    assert (inputs.RHO_L, inputs.RHO_I, inputs.CP_L, inputs.CP_I, inputs.T_0) == \
        (K["RHO_L"], K["RHO_I"], K["CP_L"], K["CP_I"], K["T_0"])
    assert inputs.LH_F0 == K["LH_F0"]
