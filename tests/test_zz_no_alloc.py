"""clm_step_noalloc against clm_step through the C ABI.

The allocation-free step runs the same kernels as clm_step with its Newton
flags and statistics in the library's static device scratch, so its states
must equal clm_step's bit for bit and its statistics must match exactly.
Two calls in a row on one GPU, the second on a smaller batch (the scratch
keeps the first call's contents), on both solvers and both models.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATS = ("max_iters_used", "n_stalled", "mean_iters", "n_nonconverged", "n_nan",
         "n_domain_error")


def _pair(name, ncol, solver, jac, jacobian=0, nsteps=3):
    """nsteps of clm_step on one context and of clm_step_noalloc on another,
    both from the same initial state and forcing; returns both states and
    per-step statistics."""
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, model_from_batch
    b = inputs.make_batch(name, ncol=ncol, nsteps=nsteps)
    w = b.workload
    out = []
    for noalloc in (False, True):
        m = model_from_batch(b, jac_precision=jac, solver=solver, jacobian=jacobian)
        sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                             tol=w.newton_tol)
        stats = []
        for _ in range(nsteps):
            st = sim.step_noalloc() if noalloc else sim.step()
            stats.append({k: getattr(st, k) for k in STATS})
        out.append((sim.state(), stats))
        m.close()
    return out


def _check(name, ncol, solver, jac, jacobian=0):
    (Ya, sa), (Yb, sb) = _pair(name, ncol, solver, jac, jacobian)
    for k in Ya:
        assert np.array_equal(Ya[k], Yb[k], equal_nan=True), (name, ncol, k)
    assert sa == sb, (name, ncol, sa, sb)
    print(f"{name} x{ncol} solver {solver} jac {jac}/{jacobian}: states equal, stats {sb[-1]}",
          flush=True)


@pytest.mark.parametrize("solver", [pytest.param(1, id="group"), pytest.param(2, id="lane")])
def test_noalloc_step_equals_step(gpu_lib, solver):
    # first call on 4,096 columns, then on a smaller batch (its scratch keeps
    # the first call's flag words beyond the smaller batch's columns)
    _check("eh_262144_uneven", 4096, solver, 64)
    _check("rre_131072_uneven", 1000, solver, 64)


def test_noalloc_newton_form_equals_step(gpu_lib):
    _check("eh_262144_uneven", 2048, 1, 32, jacobian=1)
    _check("eh_262144_uneven", 300, 1, 32, jacobian=1)


def test_noalloc_rejects_more_columns_than_the_scratch(gpu_lib):
    import ctypes as C

    from clmsoil import _lib as L
    from clmsoil.model import (ColumnBatchDomain, FreeDrainage, RichardsModel,
                               RichardsParameters, WaterFluxBC, vanGenuchten)
    n = (1 << 20) + 1
    col_ptr = np.arange(0, 2 * (n + 1), 2, dtype=np.int32)
    z_face = np.tile(np.array([-1.0, -0.5, 0.0]), n)
    one = np.ones(n)
    par = RichardsParameters(0.45 * one, vanGenuchten(2.0 * one, 1.5 * one), 1e-6 * one,
                             1e-3 * one, 0.05 * one)
    m = RichardsModel(parameters=par, domain=ColumnBatchDomain(col_ptr, z_face),
                      boundary_conditions={"top": WaterFluxBC(np.zeros(n)),
                                           "bottom": FreeDrainage()},
                      solver=L.SOLVER_GROUP)
    st = L.StepStats()
    code = m._lib.clm_step_noalloc(m._ctx, 0.0, 900.0, C.byref(st))
    assert code == L.CLM_E_ARG
    msg = m._lib.clm_last_error(m._ctx).decode()
    assert "scratch" in msg, msg
    m.close()
