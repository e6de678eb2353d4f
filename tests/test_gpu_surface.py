"""GPU: clm_soil_conductance (energy_hydrology.jl:1103-1131, 1322-1368)
against the oracle's restatement on a workload batch, after steps that dry
and freeze some top layers: 1e-13 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_soil_conductance_matches_oracle(gpu_lib, oracle_mod):
    from clmsoil import _lib as L
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, model_from_batch, soil_conductance
    b = inputs.make_batch("eh_262144_uneven", ncol=2048, nsteps=6)
    w = b.workload
    m = model_from_batch(b, solver=L.SOLVER_GROUP)
    try:
        sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                             tol=w.newton_tol)
        sim.solve(b.top_water, b.top_heat, 6, w.steps_per_slot)
        Y = sim.state()
        o = oracle_mod.OracleModel.from_batch(b)
        for d_ds, p, a in ((0.015, 1.0, 1.0), (0.05, 2.0, 0.8)):
            g, dsl = soil_conductance(m, Y, d_ds, p, a, 2.5e-5)
            go, dso = o.surface_soil_conductance(Y, d_ds, p, a, 2.5e-5)
            fin = np.isfinite(go)
            print(f"[surface] d_ds {d_ds} p {p} α {a}: {fin.sum()} finite columns, "
                  f"{(dso[fin] > 0).sum()} with a dry layer", flush=True)
            assert np.array_equal(np.isfinite(g), fin)
            assert (np.abs(g - go)[fin] <= 1e-13 * np.abs(go[fin])).all()
            assert (np.abs(dsl - dso)[fin] <= 1e-13 * np.abs(dso[fin]) + 1e-300).all()
            assert (dso[fin] > 0).sum() > 0
    finally:
        m.close()
