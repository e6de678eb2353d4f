"""The oracle against the reference's own soil-test cases (tendencies,
boundary fluxes, totals): soiltest.jl:97-647, soil_bc.jl:98-269,
conservation.jl:108-261.  The cases and their checks live in
reference_fixtures.py, shared with the GPU run of the same cases."""
import pytest

import reference_fixtures as F


@pytest.fixture
def make(oracle_mod):
    return lambda case: F.OracleBackend(oracle_mod, case)


def test_heat_only_tendency(oracle_mod, make):
    err = F.soiltest_heat_only(oracle_mod, make)
    print(f"heat only: mean |expected - dρe| / median ρe = {err:.2e}")


@pytest.mark.parametrize("ksat", [0.0, F.KSAT_DECLARED], ids=["as_run", "declared_ksat"])
@pytest.mark.parametrize("kappa_on", [False, True], ids=["water_on", "both_on"])
def test_hydrology_tendencies(oracle_mod, make, ksat, kappa_on):
    rel = F.soiltest_hydrology(oracle_mod, make, ksat, kappa_on)
    print(f"K_sat {ksat:.3g}, κ {'on' if kappa_on else 'off'}: max relative difference {rel:.2e}")
    assert rel <= 1e-5


def test_all_off_tendency(oracle_mod, make):
    F.soiltest_all_off(oracle_mod, make)


def test_moisture_state_flux(oracle_mod, make):
    F.soil_bc_moisture_state(oracle_mod, make)


def test_conductivity_at_a_state(oracle_mod, make):
    F.soil_bc_conductivity(oracle_mod, make)


def test_energy_water_free_drainage(oracle_mod, make):
    F.soil_bc_free_drainage(oracle_mod, make)


def test_richards_totals_and_flux_integral(oracle_mod, make):
    F.conservation_richards(oracle_mod, make)


def test_energy_hydrology_totals_and_flux_integrals(oracle_mod, make):
    F.conservation_energy_hydrology(oracle_mod, make)


def test_a_one_sided_face_mean_fails_the_declared_ksat_case(oracle_mod):
    """Control: the same check on a tendency built with K of the upper cell
    instead of the face mean (a plausible mistake) exceeds the 1e-5 bar."""
    import numpy as np
    zf = F.column(-1.0, 0.0, 200)
    zc = 0.5 * (zf[1:] + zf[:-1])
    vl = F.NU / 2 + F.NU / 4 * (zc + 0.5) ** 2
    thr, ks = 0.1, F.KSAT_DECLARED
    tf = 0.5 * (vl[1:] + vl[:-1])
    grad = 1.0 + F.dpsidtheta_vg(tf, thr) * F.NU / 2.0 * (0.5 * (zc[1:] + zc[:-1]) + 0.5)
    mean = 0.5 * (F.K_vg(vl[1:], thr, ks) + F.K_vg(vl[:-1], thr, ks))
    upper = F.K_vg(vl[1:], thr, ks)

    def div(q):
        fl = np.concatenate([[0.0], q, [0.0]])
        return -(fl[1:] - fl[:-1]) * 200

    a, b = div(-mean * grad), div(-upper * grad)
    assert np.abs(a - b).max() / np.abs(a).max() > 1e-3
