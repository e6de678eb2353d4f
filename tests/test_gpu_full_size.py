"""The 262,144-column EnergyHydrology workload at full size on the GPU.

The oracle cannot step all 262,144 columns in a test's time, so the full
workload is checked through properties that hold whatever the size: over a
run, each column's total water Σ (ϑ_l + θ_i ρ_i/ρ_l) Δz and total energy
Σ ρe_int Δz change by exactly what its flux integrals ∫F_water and ∫F_energy
record (the discrete operators telescope, rre.jl:502-511,
energy_hydrology.jl:1282-1320; phase change moves water between ϑ_l and θ_i
and leaves ρe_int alone), up to what the Newton stopping rule leaves in the
residual: at most the tolerance per cell and step, i.e. steps x depth x tol
in water and that times the largest ρc_s in energy.  Plus the share of
column-steps that converge, in the plain FP64, Newton-matrix and learned
forms, and the learned form's device time per step against the mixed
form's.
"""
import numpy as np
import pytest

from conftest import column_ok

pytestmark = pytest.mark.gpu

RHO_I_OVER_L = 916.7 / 1000.0
NSTEPS = 12


def _widths(b):
    dz = np.diff(b.z_face)
    keep = np.ones(dz.size, bool)
    keep[b.col_ptr[1:-1] + np.arange(b.ncol - 1)] = False  # face gaps between columns
    return dz[keep]


@pytest.mark.parametrize("form", ["fp64", "newton", "learned"])
def test_full_workload_balances(gpu_lib, form):
    from clmsoil import _lib as L
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    b = inputs.make_batch("eh_262144_uneven", nsteps=NSTEPS)
    w = b.workload
    assert b.ncol == 262144
    jac, jacobian = {"fp64": (64, L.JACOBIAN_PICARD), "newton": (32, L.JACOBIAN_NEWTON),
                     "learned": (32, L.JACOBIAN_NEWTON)}[form]
    params = None
    if form == "learned":
        from test_gpu_parity import _train_warm_start
        params, _ = _train_warm_start("eh_262144_uneven", jac, jacobian)
    m = model_from_batch(b, jac_precision=jac, solver=L.SOLVER_GROUP, jacobian=jacobian)
    if params is not None:
        ws = NewtonWarmStart(m, seed=0)
        ws.set_params(params)
        ws.enable(True)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    st, it = sim.solve(b.top_water, b.top_heat, NSTEPS, w.steps_per_slot, want_iters=True)
    Y = sim.state()
    m.close()
    dz = _widths(b)
    depth = np.add.reduceat(dz, b.col_ptr[:-1])
    Y0 = b.copy_state()

    def water(S):
        return np.add.reduceat((S["vartheta_l"] + RHO_I_OVER_L * S["theta_i"]) * dz,
                               b.col_ptr[:-1])

    def energy(S):
        return np.add.reduceat(S["rho_e_int"] * dz, b.col_ptr[:-1])

    conv = ~((it & 0x30000) != 0).any(axis=0) & column_ok(b, Y)
    dW = (water(Y) - water(Y0)) - (Y["intF_water"] - Y0["intF_water"])
    dE = (energy(Y) - energy(Y0)) - (Y["intF_energy"] - Y0["intF_energy"])
    rc_max = b.params[L.P_RHOC_DS] + 4.181e6  # ρc_s of a cell filled with liquid water
    bar_w = NSTEPS * depth * w.newton_tol
    bar_e = bar_w * rc_max
    frac_conv = 1.0 - st.n_nonconverged / (NSTEPS * b.ncol)
    print(f"{form}: {conv.sum()} of {b.ncol} columns converged in every step; column-steps "
          f"converged {frac_conv:.4f}, non-finite {st.n_nan}; water balance max "
          f"{np.abs(dW[conv]).max():.2e} m (bar {bar_w.min():.1e}..{bar_w.max():.1e}), energy "
          f"{np.abs(dE[conv]).max():.2e} J/m2 (bar {bar_e.min():.1e}..{bar_e.max():.1e}); mean "
          f"Newton iterations {st.mean_iters:.2f}", flush=True)
    assert (np.abs(dW[conv]) <= bar_w[conv]).all()
    assert (np.abs(dE[conv]) <= bar_e[conv]).all()
    assert conv.mean() >= 0.8
    assert frac_conv >= (0.9 if form == "fp64" else 0.95)


def test_learned_form_faster_than_mixed(gpu_lib):
    """All 262,144 columns, 12 steps, each form's run captured once and then
    timed as one graph replay: the learned form (warm start trained as
    bench.py trains it, on the Newton matrix) takes less device time per step
    than the mixed form (FP32 modified-Picard matrices) with at least 10%
    fewer Newton iterations, the inference included."""
    from clmsoil import _lib as L
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    from test_gpu_parity import _train_warm_start
    b = inputs.make_batch("eh_262144_uneven", nsteps=NSTEPS)
    w = b.workload
    params, _ = _train_warm_start("eh_262144_uneven", 32, L.JACOBIAN_NEWTON)
    out = {}
    for form, jacobian, ws_params in (("mixed", L.JACOBIAN_PICARD, None),
                                      ("learned", L.JACOBIAN_NEWTON, params)):
        m = model_from_batch(b, jac_precision=32, solver=L.SOLVER_GROUP, jacobian=jacobian)
        if ws_params is not None:
            ws = NewtonWarmStart(m, seed=0)
            ws.set_params(ws_params)
            ws.enable(True)
        sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                             tol=w.newton_tol)
        sim.solve(b.top_water, b.top_heat, NSTEPS, w.steps_per_slot)  # capture
        sim.reset(b.copy_state())
        st, _ = sim.solve(None, None, NSTEPS, w.steps_per_slot, upload=False)
        out[form] = (st.ms / NSTEPS, st.mean_iters)
        m.close()
    (ms_m, it_m), (ms_l, it_l) = out["mixed"], out["learned"]
    print(f"mixed {ms_m:.3f} ms/step, {it_m:.2f} iterations; learned {ms_l:.3f} ms/step, "
          f"{it_l:.2f} iterations ({100 * (1 - it_l / it_m):.0f}% fewer)", flush=True)
    assert ms_l < ms_m
    assert it_l <= 0.9 * it_m
