"""Oracle runoff for RichardsModel under RichardsAtmosDrivenFluxBC
(boundary_conditions.jl:79-212, Runoff/Runoff.jl): TOPMODELRunoff and
SurfaceRunoff.

Pinned by closed forms on designed columns (hand-derivable h∇, z∇, f_sat,
R_ss and sink values, including the saturation threshold and both branches
of max(-K_sat, P)), the reference runoff test's own column
(test/standalone/Soil/runoff.jl:95-231) on a uniform grid where the midpoint
rule makes h∇ exact, the implicit tendency's difference from a plain flux BC
(the sink and -R_ss, nothing else), and the water balance of a full step.
"""
import math

import numpy as np
import pytest

from oracle.oracle import BC_FLUX, BC_SURFACE_RUNOFF, BC_TOPMODEL, OracleModel

F_OVER, R_SB = 3.28, 1.484e-7


def column(zf, nu, thr, ksat, f_max=0.5, f_over=F_OVER, R_sb=R_SB, bc_top=BC_TOPMODEL,
           alpha=2.0, n=1.6, ss=1e-3):
    N = zf.size - 1
    rows = [nu, thr, alpha, n, ksat, ss]
    if bc_top == BC_TOPMODEL:
        rows += [f_max, f_over, R_sb]
    return OracleModel(0, BC_FLUX, np.array([0, N], dtype=np.int32), zf,
                       np.asarray(rows, dtype=np.float64).reshape(-1, 1), bc_top=bc_top)


def cache(o, vl, precip):
    Y = o.new_state()
    Y["vartheta_l"][:] = vl
    a = o.new_aux()
    o.update_cache(Y, a, np.array([precip]), None, np.array([0.0]))
    return Y, a


def test_topmodel_closed_form():
    """2 m, 20 layers of 0.1 m, θ_r = 0.05, ν = 0.45.  The bottom five hold
    ϑ = θ_r + 1.1 (ν - θ_r) (weight 1.1), the sixth exactly ν (not
    saturated: H needs twc - ν > eps), the rest 0.3.  h∇ = 0.55, z∇ = 1.45."""
    nu, thr, ks = 0.45, 0.05, 1e-6
    o = column(np.linspace(-2.0, 0.0, 21), nu, thr, ks)
    vl = np.full(20, 0.3)
    vl[:5] = thr + 1.1 * (nu - thr)
    vl[5] = nu
    _, a = cache(o, vl, -2e-6)
    zw = 1.45
    fsat = 0.5 * math.exp(-F_OVER / 2 * zw)
    rss = R_SB * math.exp(-F_OVER * zw)
    assert a["top_water"][0] == pytest.approx((1 - fsat) * -1e-6, rel=1e-13)  # P < -K_sat
    assert a["r_ss"][0] == pytest.approx(rss, rel=1e-12)
    sink = np.zeros(20)
    sink[:5] = rss / 0.55 * 1.1
    np.testing.assert_allclose(a["sink"], sink, rtol=1e-12, atol=0)
    # light rain P > -K_sat infiltrates (1 - f_sat) P
    _, a = cache(o, vl, -3e-7)
    assert a["top_water"][0] == pytest.approx((1 - fsat) * -3e-7, rel=1e-13)


def test_topmodel_unsaturated_and_oversaturated_limits():
    nu, thr, ks = 0.45, 0.05, 1e-6
    zf = np.linspace(-2.0, 0.0, 21)
    o = column(zf, nu, thr, ks, f_max=1.0)
    # no layer saturated: h∇ = 0, z∇ = depth, no sink, R_ss = R_sb e^{-f_over depth}
    _, a = cache(o, np.full(20, 0.3), -2e-6)
    assert (a["sink"] == 0).all()
    assert a["r_ss"][0] == pytest.approx(R_SB * math.exp(-F_OVER * 2.0), rel=1e-12)
    assert a["top_water"][0] == pytest.approx((1 - math.exp(-F_OVER)) * -1e-6, rel=1e-13)
    # every layer above ν (weight 1.05): h∇ = 2.1 > depth, f_sat = min(e^{+}, 1) = 1
    vl = np.full(20, thr + 1.05 * (nu - thr))
    _, a = cache(o, vl, -2e-6)
    assert a["top_water"][0] == 0.0
    rss = R_SB * math.exp(F_OVER * 0.1)
    assert a["r_ss"][0] == pytest.approx(rss, rel=1e-12)
    np.testing.assert_allclose(a["sink"], rss / 2.1 * 1.05, rtol=1e-12)


def test_reference_runoff_test_column():
    """runoff.jl:95-231 on 15 uniform layers: ϑ_l = 0.6 - 0.006 (z + 50),
    ν = 0.5, θ_r = 0, K_sat = 1e-6, f_max = 0.5.  Cells 0..4 (centres up to
    z + 50 = 15) exceed ν; the midpoint rule integrates the linear ϑ/ν
    exactly: h∇ = ∫_0^{50/3} (1.2 - 0.012 s) ds = 55/3."""
    o = column(np.linspace(-50.0, 0.0, 16), 0.5, 0.0, 1e-6)
    zc = -50.0 + (np.arange(15) + 0.5) * 50.0 / 15
    vl = 0.6 - 0.006 * (zc + 50.0)
    for precip in (-1.5e-6, -5e-7):  # the test's P range, -1e-6 ± 5e-7
        _, a = cache(o, vl, precip)
        zw = 50.0 - 55.0 / 3.0
        assert a["r_ss"][0] == pytest.approx(R_SB * math.exp(-F_OVER * zw), rel=1e-10)
        fsat = 0.5 * math.exp(-F_OVER / 2 * zw)
        assert a["top_water"][0] == pytest.approx((1 - fsat) * max(-1e-6, precip), rel=1e-13)
        assert (a["sink"][:5] > 0).all() and (a["sink"][5:] == 0).all()
        dz = 50.0 / 15
        assert (a["sink"] * dz).sum() == pytest.approx(a["r_ss"][0], rel=1e-12)


def test_imp_tendency_is_flux_tendency_minus_sink():
    """compute_imp_tendency with the implicit TOPMODELSubsurfaceRunoff source
    (rre.jl:195-200, Runoff.jl:321-332) differs from the same column under a
    prescribed flux equal to the infiltration by exactly -sink per layer and
    -R_ss in ∫F."""
    nu, thr, ks = 0.45, 0.05, 1e-6
    zf = -2.0 * (1.05 ** np.arange(20, -1, -1) - 1) / (1.05 ** 20 - 1)  # stretched
    o = column(zf, nu, thr, ks)
    vl = np.linspace(0.46, 0.25, 20)
    Y, a = cache(o, vl, -2e-6)
    assert (a["sink"] > 0).any() and (a["sink"] == 0).any()
    d = o.imp_tendency(Y, a)
    of = column(zf, nu, thr, ks, bc_top=BC_FLUX)
    Yf, af = cache(of, vl, a["top_water"][0])
    df = of.imp_tendency(Yf, af)
    np.testing.assert_allclose(d["vartheta_l"], df["vartheta_l"] - a["sink"], rtol=1e-14,
                               atol=1e-25)
    assert d["intF_water"][0] == pytest.approx(df["intF_water"][0] - a["r_ss"][0], rel=1e-14)


@pytest.mark.parametrize("wt", [0.0, 0.7])
def test_step_water_balance(wt):
    """One ARS111 step with the runoff lagged at u_n and a no-flux bottom:
    Σ Δz_c Δϑ = -Δt (infiltration + R_ss) when a layer is saturated, and
    -Δt infiltration (the sink is 0, ∫F still takes R_ss) when none is."""
    nu, thr, ks = 0.45, 0.05, 1e-5
    zf = np.linspace(-2.0, 0.0, 31)
    zc = 0.5 * (zf[1:] + zf[:-1])
    o = column(zf, nu, thr, ks, R_sb=1e-6)
    zwt = -2.0 + wt * 2.0
    vl = np.where(zc < zwt, nu + 1e-3 * (zwt - zc), 0.3)
    Y, a = cache(o, vl, -2e-5)
    infil, rss = a["top_water"][0], a["r_ss"][0]
    dt = 1800.0
    it, _ = o.step(Y, dt, 50, 1e-12, np.array([-2e-5]), None, np.array([0.0]))
    assert it[0] < 50
    dw = ((Y["vartheta_l"] - vl) * np.diff(zf)).sum()
    assert Y["intF_water"][0] == pytest.approx(-dt * (infil + rss), rel=1e-12)
    expect = -dt * (infil + rss) if wt > 0 else -dt * infil
    assert dw == pytest.approx(expect, rel=1e-7, abs=1e-12)
    if wt > 0:
        assert rss > 0 and (a["sink"] > 0).any()


def test_surface_runoff():
    """SurfaceRunoff (Runoff.jl:107-148): no infiltration when the top layer
    exceeds ν, else max(-K_sat, P); no subsurface sink."""
    nu, thr, ks = 0.45, 0.05, 1e-6
    o = column(np.linspace(-1.0, 0.0, 11), nu, thr, ks, bc_top=BC_SURFACE_RUNOFF)
    vl = np.full(10, 0.3)
    for precip, expect in ((-2e-6, -1e-6), (-4e-7, -4e-7)):
        _, a = cache(o, vl, precip)
        assert a["top_water"][0] == expect
    vl[-1] = nu  # at ν: not saturated
    assert cache(o, vl, -2e-6)[1]["top_water"][0] == -1e-6
    vl[-1] = nu + 1e-4
    assert cache(o, vl, -2e-6)[1]["top_water"][0] == 0.0
    vl[-1], vl[0] = 0.3, nu + 1e-3  # only the top layer counts
    assert cache(o, vl, -2e-6)[1]["top_water"][0] == -1e-6
