"""Paired waves: the EnergyHydrology step with columns of 33-40 and 17-24
layers sharing waves (clm_set_pairing) against the same step with each bin
on its own shape.

The paired kernel gives each column only its own rows, padding, reductions
and broadcasts, so every column's state and Newton flags must equal the
unpaired run's bit for bit: in the FP64, mixed, Newton-matrix and
warm-started forms, with the deep bin larger than the shallow one (waves
with no second column) and smaller (shallow columns left to their own
shape), over graph-replayed runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = ((16, 16, 1), (24, 32, 1), (32, 32, 1), (40, 64, 1))


def _subset(b, deep_share):
    """Columns of every bucket, with the <= 24 bucket cut to `deep_share`
    times the <= 40 bucket's size (or the other way round)."""
    n = b.nlayers
    deep = np.nonzero((n > 32) & (n <= 40))[0]
    shallow = np.nonzero((n > 16) & (n <= 24))[0]
    rest = np.nonzero((n <= 16) | ((n > 24) & (n <= 32)))[0]
    if deep_share >= 1.0:
        shallow = shallow[: int(len(deep) / deep_share)]
    else:
        deep = deep[: int(len(shallow) * deep_share)]
    cols = np.sort(np.concatenate([deep, shallow, rest]))
    return b.column_subset(cols), len(deep), len(shallow)


def _run(b, jac, jacobian, pairing, ws_params=None, nsteps=6, stream=None):
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    w = b.workload
    m = model_from_batch(b, jac_precision=jac, solver=1, jacobian=jacobian, stream=stream)
    for mx, g, r in SHAPES:
        m.set_shape(mx, g, r)
    m.set_pairing(pairing)
    active = m.pairing_active()
    if ws_params is not None:
        ws = NewtonWarmStart(m, seed=0)
        ws.set_params(ws_params)
        ws.enable(True)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    st, it = sim.solve(b.top_water, b.top_heat, nsteps, w.steps_per_slot, want_iters=True)
    Y = sim.state()
    m.close()
    return Y, it, st, active


@pytest.mark.parametrize("deep_share", [1.5, 0.6], ids=["deep_bin_larger", "shallow_bin_larger"])
@pytest.mark.parametrize("form", ["fp64", "mixed", "newton", "warm"])
def test_paired_waves_equal_separate_launches(gpu_lib, form, deep_share):
    from clmsoil import _lib as L
    from clmsoil import inputs
    from clmsoil.model import NewtonWarmStart, model_from_batch
    b0 = inputs.make_batch("eh_262144_uneven", ncol=6000, nsteps=6)
    b, nd, ns = _subset(b0, deep_share)
    jac, jacobian = {"fp64": (64, L.JACOBIAN_PICARD), "mixed": (32, L.JACOBIAN_PICARD),
                     "newton": (32, L.JACOBIAN_NEWTON), "warm": (32, L.JACOBIAN_NEWTON)}[form]
    params = None
    if form == "warm":  # a random network: the guesses differ by column
        m = model_from_batch(b, jac_precision=jac, solver=1, jacobian=jacobian)
        ws = NewtonWarmStart(m, seed=3)
        rng = np.random.default_rng(11)
        params = (ws.params() + 0.02 * rng.standard_normal(ws.n_params)).astype(np.float32)
        m.close()
    Yp, itp, stp, act_p = _run(b, jac, jacobian, True, params)
    Ys, its, sts, act_s = _run(b, jac, jacobian, False, params)
    assert act_p and not act_s
    for k in Yp:
        assert np.array_equal(Yp[k], Ys[k], equal_nan=True), (form, k)
    assert np.array_equal(itp, its), form
    print(f"{form}: {b.ncol} columns ({nd} of 33-40 layers, {ns} of 17-24): paired run equals "
          f"separate launches bit for bit; mean Newton iterations {stp.mean_iters:.2f}, "
          f"{stp.ms / 6:.3f} vs {sts.ms / 6:.3f} ms/step", flush=True)


def test_callers_stream_equals_own_stream(gpu_lib):
    """A context created on a caller's HIP stream against one on the
    library's own stream: a step's depth bins fork from and join back into
    whichever stream the context has, in the captured run, so the states and
    Newton counts are equal bit for bit.  Learned form on the Newton matrix
    with paired waves and a perturbed network.  The caller's stream comes
    from the HIP runtime the library itself resolved (non-blocking, default
    priority, as torch creates its streams)."""
    import ctypes as C
    from clmsoil import _lib as L
    from clmsoil import inputs
    from clmsoil.model import NewtonWarmStart, model_from_batch
    b0 = inputs.make_batch("eh_262144_uneven", ncol=6000, nsteps=8)
    b, _, _ = _subset(b0, 1.0)
    m = model_from_batch(b, jac_precision=32, solver=1, jacobian=L.JACOBIAN_NEWTON)
    ws = NewtonWarmStart(m, seed=3)
    rng = np.random.default_rng(13)
    params = (ws.params() + 0.02 * rng.standard_normal(ws.n_params)).astype(np.float32)
    m.close()
    hip = C.CDLL("libamdhip64.so.7")  # the copy already in the process, by its soname
    hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    s = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(s), 1) == 0 and s.value  # non-blocking
    try:
        Yc, itc, stc, act_c = _run(b, 32, L.JACOBIAN_NEWTON, True, params, 8, s.value)
    finally:
        hip.hipStreamDestroy(s)
    Yo, ito, sto, act_o = _run(b, 32, L.JACOBIAN_NEWTON, True, params, 8)
    assert act_c and act_o
    for k in Yc:
        assert np.array_equal(Yc[k], Yo[k], equal_nan=True), k
    assert np.array_equal(itc, ito)
    print(f"{b.ncol} columns, 8 steps on the caller's stream equal the library's own stream bit "
          f"for bit; mean Newton iterations {stc.mean_iters:.2f}", flush=True)
