"""Checkpoint / restart (checkpoints.jl:125-195; cf. test/integrated/restart.jl):
a run restarted from a checkpoint continues bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["rre_131072_uneven", "eh_262144_uneven"])
def test_restart_continues_bit_for_bit(gpu_lib, tmp_path, name):
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, model_from_batch
    b = inputs.make_batch(name, ncol=512, nsteps=12)
    w = b.workload
    m = model_from_batch(b, solver=1)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    sim.solve(b.top_water, b.top_heat, 6, w.steps_per_slot)
    path = sim.save_checkpoint(str(tmp_path))
    t_ck = 6 * w.dt
    assert os.path.basename(path) == f"day{int(t_ck // 86400)}.{int(t_ck % 86400)}.clmchk"
    # continue 6 more steps (forcing slots of steps 6..11)
    tail_w = b.top_water[1:]
    tail_h = None if b.top_heat is None else b.top_heat[1:]
    sim.solve(tail_w, tail_h, 6, w.steps_per_slot)
    Ya = sim.state()
    m.close()
    m2 = model_from_batch(b, solver=1)
    s2 = LandSimulation(m2, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    assert s2.restore_checkpoint(path)
    assert s2.t == t_ck
    s2.solve(tail_w, tail_h, 6, w.steps_per_slot)
    Yb = s2.state()
    for k in Ya:
        assert np.array_equal(Ya[k], Yb[k], equal_nan=True), k
    m2.close()
    # a model with other parameters reads the file but reports the difference
    b3 = inputs.make_batch(name, ncol=512)
    b3.params[0] *= 1.01
    m3 = model_from_batch(b3, solver=1)
    s3 = LandSimulation(m3, b3.copy_state(), w.dt)
    assert not s3.restore_checkpoint(path)
    m3.close()
    # a different geometry is rejected
    from clmsoil._lib import ClmError
    b4 = inputs.make_batch(name, ncol=256)
    m4 = model_from_batch(b4, solver=1)
    s4 = LandSimulation(m4, b4.copy_state(), w.dt)
    with pytest.raises(ClmError):
        s4.restore_checkpoint(path)
    m4.close()
