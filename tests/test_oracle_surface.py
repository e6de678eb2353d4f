"""The surface vapour-conductance closures of EnergyHydrology
(energy_hydrology.jl:1103-1131, 1322-1368) in the oracle, pinned by the
reference's case in test/standalone/Soil/climate_drivers.jl:265-289 and by
closed forms."""
import numpy as np

from clmsoil import inputs

EPS = np.finfo(np.float64).eps
# climate_drivers.jl:32-40 (ν, θ_r, van Genuchten α, n); d_ds, evap_p and
# evap_α as soiltest.jl:185-212 passes them to the direct constructor
NU, THR, VG_N = 0.495, 0.1, 2.0
D_DS, EVAP_P, EVAP_A = 0.015, 1.0, 1.0
D_VAPOR = 2.5e-5  # m2/s; the closures' output scales with it


def test_vg_sc_closed_form(oracle_mod):
    L = oracle_mod.lib()
    # n = 2: m = 1/2, ((n-1)/n)^(1-2n) = (1/2)^-3 = 8, S_c = 9^(-1/2) = 1/3
    assert abs(L.orc_vg_sc(2.0) - 1.0 / 3.0) <= 2 * EPS
    for n in (1.3, 1.8, 2.7):
        m = 1 - 1 / n
        assert abs(L.orc_vg_sc(n) - (1 + ((n - 1) / n) ** (1 - 2 * n)) ** -m) <= 4 * EPS


def test_climate_drivers_dry_layer_and_conductance(oracle_mod):
    """climate_drivers.jl:265-289: over θ = range(θ_r + eps, ν, 5) the dry
    layer's smallest value is 0 and its largest ((S_c - S_1)/S_1)^p d_ds; the
    conductance grows from the driest surface to the next."""
    L = oracle_mod.lib()
    theta = np.linspace(THR + EPS, NU, 5)
    S = np.array([L.orc_effective_saturation(NU, t, THR) for t in theta])
    Sc = L.orc_vg_sc(VG_N) * EVAP_A
    dsl = np.array([L.orc_dry_soil_layer_thickness(s, Sc, D_DS, EVAP_P) for s in S])
    assert (dsl / D_DS).min() == 0.0
    assert (dsl / D_DS).max() == ((Sc - S[0]) / S[0]) ** EVAP_P
    g = np.array([L.orc_soil_conductance(s, L.orc_vg_sc(VG_N), D_DS, EVAP_P, EVAP_A, D_VAPOR)
                  for s in S])
    assert g[0] < g[1]
    # above S_c there is no dry layer and the conductance is D_vapor / eps
    assert g[-1] == D_VAPOR / EPS
    # α S_c is clamped to [0, 1]
    assert L.orc_dry_soil_layer_thickness(0.5, 3.0, D_DS, 1.0) == D_DS * (1.0 - 0.5) / 0.5


def test_surface_extrapolation_of_a_linear_profile(oracle_mod):
    """A θ_l linear in depth extrapolates to its exact surface value
    (linear_interpolation_to_surface!), so the conductance matches the
    closure evaluated there; a surface value below θ_r is floored."""
    b = inputs.make_batch("eh_262144_uneven", ncol=64)
    o = oracle_mod.OracleModel.from_batch(b)
    Y = b.copy_state()
    cp = b.col_ptr
    L = oracle_mod.lib()
    want = np.zeros(b.ncol)
    for c in range(b.ncol):
        a, e = cp[c], cp[c + 1]
        zf = b.z_face[a + c:e + c + 1]
        zc = 0.5 * (zf[1:] + zf[:-1])
        nu, thr = b.params[0, c], b.params[1, c]
        profile = thr + 0.2 * (nu - thr) + 0.01 * c / b.ncol * (zf[-1] - zc)  # linear in z
        Y["vartheta_l"][a:e] = profile
        Y["theta_i"][a:e] = 0.0
        t0 = thr + 0.2 * (nu - thr)  # value at the surface z = 0 (zf[-1] = 0)
        S = L.orc_effective_saturation(nu, t0, thr)
        want[c] = L.orc_soil_conductance(S, L.orc_vg_sc(b.params[3, c]), D_DS, EVAP_P, EVAP_A,
                                         D_VAPOR)
    g, _ = o.surface_soil_conductance(Y, D_DS, EVAP_P, EVAP_A, D_VAPOR)
    assert np.allclose(g, want, rtol=1e-12, atol=0)
    # a steep drying profile whose extrapolation falls below θ_r: floored
    c = 0
    a, e = cp[c], cp[c + 1]
    thr = b.params[1, c]
    Y["vartheta_l"][a:e] = thr + 0.05
    Y["vartheta_l"][e - 1] = thr + 1e-4
    g2, dsl = o.surface_soil_conductance(Y, D_DS, EVAP_P, EVAP_A, D_VAPOR)
    S = L.orc_effective_saturation(b.params[0, c], thr + EPS, thr)
    Sc = L.orc_vg_sc(b.params[3, c])
    assert dsl[c] == L.orc_dry_soil_layer_thickness(S, Sc, D_DS, EVAP_P)
