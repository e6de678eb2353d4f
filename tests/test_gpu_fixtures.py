"""The HIP library's hooks and column totals on the reference's own soil-test
cases (soiltest.jl:97-647, soil_bc.jl:98-269, conservation.jl:108-261), with
the same checks as the oracle's run of them (reference_fixtures.py), and the
device totals (clm_totals) against the oracle's on a workload batch."""
import numpy as np
import pytest

import reference_fixtures as F
from clmsoil import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def make(gpu_lib):
    made = []

    def mk(case):
        b = F.GpuBackend(case)
        made.append(b)
        return b
    yield mk
    for b in made:
        b.close()


def test_heat_only_tendency(gpu_lib, oracle_mod, make):
    print("[fixtures] soiltest heat only", flush=True)
    err = F.soiltest_heat_only(oracle_mod, make)
    print(f"heat only: mean |expected - dρe| / median ρe = {err:.2e}", flush=True)


@pytest.mark.parametrize("ksat", [0.0, F.KSAT_DECLARED], ids=["as_run", "declared_ksat"])
@pytest.mark.parametrize("kappa_on", [False, True], ids=["water_on", "both_on"])
def test_hydrology_tendencies(gpu_lib, oracle_mod, make, ksat, kappa_on):
    print(f"[fixtures] soiltest K_sat {ksat:.3g} kappa {kappa_on}", flush=True)
    rel = F.soiltest_hydrology(oracle_mod, make, ksat, kappa_on)
    print(f"K_sat {ksat:.3g}, κ {'on' if kappa_on else 'off'}: max relative difference "
          f"{rel:.2e}", flush=True)
    assert rel <= 1e-5


def test_all_off_tendency(gpu_lib, oracle_mod, make):
    F.soiltest_all_off(oracle_mod, make)


def test_moisture_state_flux(gpu_lib, oracle_mod, make):
    F.soil_bc_moisture_state(oracle_mod, make)


def test_conductivity_at_a_state(gpu_lib, oracle_mod, make):
    F.soil_bc_conductivity(oracle_mod, make)


def test_energy_water_free_drainage(gpu_lib, oracle_mod, make):
    F.soil_bc_free_drainage(oracle_mod, make)


def test_richards_totals_and_flux_integral(gpu_lib, oracle_mod, make):
    F.conservation_richards(oracle_mod, make)


def test_energy_hydrology_totals_and_flux_integrals(gpu_lib, oracle_mod, make):
    F.conservation_energy_hydrology(oracle_mod, make)


@pytest.mark.parametrize("name", ["rre_131072_uneven", "eh_262144_uneven"])
@pytest.mark.parametrize("solver", [L.SOLVER_GROUP, L.SOLVER_LANE], ids=["group", "lane"])
def test_device_totals_match_oracle(gpu_lib, oracle_mod, name, solver):
    """clm_totals on 3,000 columns after two steps (a state with ice for
    EnergyHydrology) against orc_totals: 1e-13 relative (energy: of Σ |ρe| Δz)."""
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, model_from_batch
    b = inputs.make_batch(name, ncol=3000)
    w = b.workload
    m = model_from_batch(b, solver=solver)
    try:
        sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                             tol=w.newton_tol)
        sim.solve(b.top_water[:1], None if b.top_heat is None else b.top_heat[:1], 2, 6)
        wg, eg = sim.totals()
        Y = sim.state()
        o = oracle_mod.OracleModel.from_batch(b)
        wo, eo = o.totals(Y)
        fin = np.isfinite(wo)
        print(f"[fixtures] totals {name} solver {solver}: {fin.sum()} finite columns", flush=True)
        assert np.array_equal(np.isfinite(wg), fin)
        assert (np.abs(wg - wo)[fin] <= 1e-13 * np.abs(wo[fin])).all()
        if eo is not None:
            # ρe_int changes sign across a column (ρc_s (T - T_0) - θ_i ρ_i L_f),
            # so Σ ρe Δz can cancel: its rounding scales with Σ |ρe| Δz
            fe = np.isfinite(eo)
            dz = np.diff(b.z_face)[np.arange(b.ncell) + np.repeat(np.arange(b.ncol), b.nlayers)]
            scale = np.add.reduceat(np.abs(Y["rho_e_int"]) * dz, b.col_ptr[:-1])
            err = np.abs(eg - eo)[fe] / scale[fe]
            print(f"[fixtures] energy totals: max error {err.max():.2e} of sum |ρe| Δz",
                  flush=True)
            assert (err <= 1e-13).all()
    finally:
        m.close()
