"""GPU checks of the learned Newton warm start (clm_ws_* through the C ABI).

Bars:
  * features: equal to the CPU restatement rounded to BF16 on >= 99.5% of
    entries, every entry within 1 BF16 ulp (2^-7 relative) or 1e-6;
  * network forward / loss / gradient vs the CPU restatement with the same
    BF16 rounding points: relative L2 error <= 1e-2 per output and per
    layer gradient (rounding of intermediates may flip on near-ties);
  * warm start with a zero guess (x0 = U0), and with the nonzero guesses of
    perturbed parameters: the oracle's warm-started step from the same
    guesses to 1e-10 on well-conditioned columns;
  * trained warm start: same root as plain Newton to the Newton tolerance,
    and it reports the iterations it saves.
"""
import numpy as np
import pytest

from conftest import column_ok, rel_err

pytestmark = pytest.mark.gpu


def _model(batch, jac=64):
    from clmsoil.model import model_from_batch
    return model_from_batch(batch, jac_precision=jac, solver=1)


def _set_forcing(m, b, slot=0):
    from clmsoil.model import HeatFluxBC, WaterFluxBC, WaterHeatBC
    if b.top_heat is None:
        m.boundary_conditions["top"] = WaterFluxBC(b.top_water[slot])
    else:
        m.boundary_conditions["top"] = WaterHeatBC(WaterFluxBC(b.top_water[slot]),
                                                   HeatFluxBC(b.top_heat[slot]))
    m.set_forcing()


@pytest.mark.parametrize("name", ["rre_131072_uneven", "eh_262144_uneven"])
def test_features_match_cpu(gpu_lib, name):
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart
    from oracle import mlp_oracle as mo
    from oracle.oracle import constants
    b = inputs.make_batch(name, ncol=300)
    m = _model(b)
    _set_forcing(m, b)
    sim = LandSimulation(m, b.copy_state(), b.workload.dt)
    sim._push()
    ws = NewtonWarmStart(m, seed=1)
    X = ws.features(b.workload.dt)
    ref = mo.features(b, b.copy_state(), b.top_water[0], None if b.top_heat is None else
                      b.top_heat[0], b.workload.dt, constants())
    refb = mo.to_bf16(ref.astype(np.float32))
    same = X == refb
    print(f"{name}: features equal on {same.mean():.5f} of entries", flush=True)
    assert same.mean() >= 0.995
    assert (np.abs(X - refb) <= np.maximum(2.0 ** -7 * np.abs(refb), 1e-6)).all()
    m.close()


@pytest.mark.parametrize("eh,M", [(False, 256), (True, 256), (True, 2048)])
def test_network_forward_and_gradient_match_cpu(gpu_lib, eh, M):
    """M = 2048 takes the split-K path of the weight-gradient GEMMs."""
    from clmsoil import inputs
    from clmsoil.model import NewtonWarmStart
    from oracle import mlp_oracle as mo
    b = inputs.make_batch("eh_262144_uneven" if eh else "rre_131072_uneven", ncol=64)
    m = _model(b)
    ws = NewtonWarmStart(m, seed=7)
    F, Fp, H, No = mo.dims(eh)
    assert (ws.F, ws.Fp, ws.H, ws.Nout) == (F, Fp, H, No)
    rng = np.random.default_rng(11)
    X = np.zeros((M, Fp), dtype=np.float32)
    X[:, :F] = rng.standard_normal((M, F))
    T = rng.standard_normal((M, No)).astype(np.float32)
    mask = (rng.random(M) > 0.25).astype(np.float32)
    p = ws.params()
    # perturb the output layer so that every layer carries a sizeable gradient
    p = (p + 0.02 * rng.standard_normal(p.size)).astype(np.float32)
    ws.set_params(p)
    Y = ws.forward(X)
    Yr, _ = mo.forward(p, eh, X)
    ey = np.linalg.norm(Y - Yr) / np.linalg.norm(Yr)
    loss, g = ws.gradient(X, T, mask)
    lr, gr = mo.loss_and_grad(p, eh, X, T, mask.astype(np.float64))
    print(f"eh={eh}: forward rel err {ey:.2e}; loss {loss:.6g} vs {lr:.6g}", flush=True)
    assert ey <= 1e-2
    assert abs(loss - lr) <= 1e-2 * abs(lr)
    o = 0
    for li, (n, k) in enumerate(mo.layer_shapes(eh)):
        for part, size in (("W", n * k), ("b", n)):
            a, r = g[o:o + size], gr[o:o + size]
            e = np.linalg.norm(a - r) / np.linalg.norm(r)
            print(f"  layer {li} {part}: grad rel err {e:.2e} (|g| {np.linalg.norm(r):.3g})",
                  flush=True)
            assert e <= 1e-2, (li, part, e)
            o += size
    assert o == p.size
    m.close()


@pytest.mark.parametrize("name", ["rre_131072_uneven", "eh_262144_uneven"])
def test_fused_inference_equals_layer_path(gpu_lib, name):
    """The one-kernel inference (features, four layers and guesses with the
    activations in LDS) against the layer-by-layer kernels: the same
    accumulation order, so the network outputs and the ϑ_l guesses are equal
    bit for bit (the ρe_int guesses to an FP64 ulp); and against the CPU
    restatement (features -> BF16 network -> node-to-cell resampling) to the
    BF16 bar.  300 columns: the last 128-column block is partial."""
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart
    from oracle import mlp_oracle as mo
    from oracle.oracle import constants
    b = inputs.make_batch(name, ncol=300)
    eh = b.top_heat is not None
    m = _model(b)
    _set_forcing(m, b)
    sim = LandSimulation(m, b.copy_state(), b.workload.dt)
    sim._push()
    ws = NewtonWarmStart(m, seed=2)
    rng = np.random.default_rng(4)
    p = (ws.params() + 0.05 * rng.standard_normal(ws.n_params)).astype(np.float32)
    ws.set_params(p)
    dt = b.workload.dt
    gv1, ge1 = ws.guess(dt, fused=True)
    gv0, ge0 = ws.guess(dt, fused=False)
    assert np.isfinite(gv1).all() and np.abs(gv1).max() > 0
    assert (gv1 == gv0).all()
    if eh:
        # the node-to-cell interpolation and ρc_s(u) are FP64 expressions that
        # the two kernels may contract into FMAs differently: rounding level
        # against the guesses' scale (interpolation can cancel towards zero)
        e = np.abs(ge1 - ge0).max() / np.abs(ge0).max()
        print(f"{name}: energy guess, fused vs layer path max err {e:.2e} of max |guess|",
              flush=True)
        assert e <= 1e-14
    X = mo.features(b, b.copy_state(), b.top_water[0], b.top_heat[0] if eh else None, dt,
                    constants())
    Y, _ = mo.forward(p, eh, mo.to_bf16(X.astype(np.float32)))
    ref = np.concatenate([mo.at_cells(Y[c, :32], n) * 1e-2 for c, n in enumerate(b.nlayers)])
    e = np.linalg.norm(gv1 - ref) / np.linalg.norm(ref)
    print(f"{name}: fused guess vs CPU restatement rel L2 {e:.2e}", flush=True)
    assert e <= 1e-2
    m.close()


def _one_step(b, sim, m):
    w = b.workload
    th = None if b.top_heat is None else b.top_heat[:1]
    _, it = sim.solve(b.top_water[:1], th, 1, 1, want_iters=True)
    return sim.state(), it[0]


@pytest.mark.parametrize("params", ["zero", "perturbed"])
@pytest.mark.parametrize("name", ["rre_131072_uneven", "eh_262144_uneven"])
def test_zero_guess_matches_oracle_warm_start(gpu_lib, name, params):
    """The step kernels' warm-start path against the oracle's orc_step_guess
    given the same guesses.  With all network parameters zero the guess is
    exactly 0.  With perturbed parameters the guesses are nonzero and differ
    by cell; the oracle gets those of clm_ws_guess (k_guess over the network
    rows), while the step kernels form theirs from the same rows themselves:
    a guess formed differently, or a row missed or stale, would change the
    iterations."""
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart
    from oracle.oracle import OracleModel
    b = inputs.make_batch(name, ncol=512)
    w = b.workload
    m = _model(b)
    ws = NewtonWarmStart(m, seed=3)
    if params == "zero":
        ws.set_params(np.zeros(ws.n_params, dtype=np.float32))
    else:
        rng = np.random.default_rng(5)
        ws.set_params((ws.params() + 0.01 * rng.standard_normal(ws.n_params)).astype(np.float32))
    ws.enable(True)
    sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters, tol=w.newton_tol)
    if params == "zero":
        gv = np.zeros(b.ncell)
        ge = gv if b.top_heat is not None else None
    else:
        _set_forcing(m, b)  # the forcing of the step's slot 0
        sim._push()
        gv, ge = ws.guess(w.dt, fused=True)
        assert np.abs(gv).max() > 0 and np.unique(gv).size > b.ncol
    Yg, itg = _one_step(b, sim, m)
    fw = b.top_water[0].copy()
    fh = None if b.top_heat is None else b.top_heat[0].copy()
    outs = {}
    for v in ("", "xprec", "fma"):
        o = OracleModel.from_batch(b, variant=v)
        Y = b.copy_state()
        it, _ = o.step(Y, w.dt, w.newton_max_iters, w.newton_tol, fw, fh, guess_v=gv,
                       guess_e=ge)
        outs[v] = (Y, it)
    Yo, ito = outs[""]
    ok = column_ok(b, Yo) & column_ok(b, Yg) & (ito < w.newton_max_iters)
    sens = np.zeros(b.ncol)
    keys = ["vartheta_l"] + (["rho_e_int"] if b.top_heat is not None else [])
    for v in ("xprec", "fma"):
        Yv, itv = outs[v]
        ok &= column_ok(b, Yv) & (itv == ito)
        for k in keys:
            fl = 1.0 if k == "rho_e_int" else 1e-12
            sens = np.maximum(sens, np.maximum.reduceat(
                np.where(np.repeat(ok, b.nlayers), rel_err(Yv[k], Yo[k], fl), 0.0),
                b.col_ptr[:-1]))
    well = ok & (sens <= 1e-12) & ((itg & 0xFFFF) == ito)
    assert well.mean() > 0.6, well.mean()
    assert ((itg & 0xFFFF) == ito)[ok].mean() >= 0.98
    assert not (itg & 0x80000).any()
    for k in keys:
        fl = 1.0 if k == "rho_e_int" else 1e-12
        e = np.where(np.repeat(well, b.nlayers), rel_err(Yg[k], Yo[k], fl), 0.0)
        print(f"{name} {k}: {params} warm start max rel err {e.max():.2e} on {well.sum()} "
              f"columns", flush=True)
        assert e.max() <= 1e-10
    m.close()


@pytest.mark.parametrize("name,jac", [("eh_262144_uneven", 32), ("rre_131072_uneven", 64)])
def test_trained_warm_start_saves_iterations(gpu_lib, name, jac):
    """Train on columns from other seeds, evaluate on the workload's own
    columns: same converged states as plain Newton, fewer iterations."""
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart
    tname = inputs.training_workload(name)
    bt = inputs.make_batch(tname, ncol=8192, nsteps=12)
    w = bt.workload
    m = _model(bt, jac)
    ws = NewtonWarmStart(m, seed=5)
    sim = LandSimulation(m, bt.copy_state(), w.dt, max_iters=w.newton_max_iters,
                         tol=w.newton_tol)
    rows = ws.collect(sim, bt.top_water, bt.top_heat, 12, w.steps_per_slot)
    assert rows == 12 * 8192
    losses = ws.train(epochs=8, batch=2048, lr=1e-3, seed=9)
    print(f"{name}: training loss by epoch {np.array2string(losses, precision=4)}", flush=True)
    assert np.isfinite(losses).all() and losses[-1] < 0.7 * losses[0]
    params = ws.params()
    m.close()
    # evaluation on the workload's own first columns, steps 0..5
    b = inputs.make_batch(name, ncol=4096, nsteps=6)
    res = {}
    for on in (False, True):
        m2 = _model(b, jac)
        ws2 = NewtonWarmStart(m2, seed=0)
        ws2.set_params(params)
        ws2.enable(on)
        s2 = LandSimulation(m2, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                            tol=w.newton_tol)
        st, it = s2.solve(b.top_water, b.top_heat, 6, w.steps_per_slot, want_iters=True)
        res[on] = (s2.state(), it)
        m2.close()
    (Yp, itp), (Yw, itw) = res[False], res[True]
    good = ~(((itp | itw) & 0x70000) != 0).any(axis=0)
    good &= column_ok(b, Yp) & column_ok(b, Yw)
    kp, kw = (itp & 0xFFFF)[:, good], (itw & 0xFFFF)[:, good]
    fb = ((itw & 0x80000) != 0).mean()
    dcol = np.maximum.reduceat(np.abs(Yw["vartheta_l"] - Yp["vartheta_l"]), b.col_ptr[:-1])[good]
    q50, q99 = np.quantile(dcol, [0.5, 0.99])
    print(f"{name}: mean Newton iterations plain {kp.mean():.3f} -> warm start {kw.mean():.3f} "
          f"({100 * (1 - kw.mean() / kp.mean()):.1f}% fewer) on {good.sum()} columns x 6 steps; "
          f"fallbacks {fb:.4f}; per-column max |Δϑ_l| vs plain: median {q50:.1e}, 99% {q99:.1e}, "
          f"max {dcol.max():.1e}", flush=True)
    assert good.mean() > 0.5
    # two Newton solutions to 1e-8 from different starts agree to about the
    # tolerance; over six steps a few slowly contracting (modified-Picard)
    # columns separate further, as they do under a 1-ulp change of the input
    assert q50 <= 1e-7 and q99 <= 1e-5
    # a warm start that leads Newton into a stall (increments below tol far
    # from a root) falls back to U0, so no column lands far from plain Newton's
    assert dcol.max() <= 1e-4
    assert kw.mean() <= kp.mean()


def test_run_with_warm_start_equals_steps(gpu_lib):
    """An eight-step run captured as one graph against the same eight steps
    one clm_step call at a time: the states are equal bit for bit and the
    Newton iterations agree.  The learned form on the Newton matrix, paired
    waves active, a perturbed network so the guesses differ by column, and
    the forcing slot changing inside the run (six steps per slot), so each
    step's feature rows must use that step's state and forcing."""
    from clmsoil import _lib as L
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    nsteps = 8
    b = inputs.make_batch("eh_262144_uneven", ncol=3000, nsteps=nsteps)
    w = b.workload
    assert len({b.forcing_slot(s) for s in range(nsteps)}) > 1
    params = None
    outs = []
    for mode in ("run", "steps"):
        m = model_from_batch(b, jac_precision=32, solver=L.SOLVER_GROUP,
                             jacobian=L.JACOBIAN_NEWTON)
        for mx, g, r in ((16, 16, 1), (24, 32, 1), (32, 32, 1), (40, 64, 1)):
            m.set_shape(mx, g, r)
        assert m.pairing_active()
        ws = NewtonWarmStart(m, seed=3)
        if params is None:
            rng = np.random.default_rng(7)
            params = (ws.params() + 0.02 * rng.standard_normal(ws.n_params)).astype(np.float32)
        ws.set_params(params)
        ws.enable(True)
        sim = LandSimulation(m, b.copy_state(), w.dt, max_iters=w.newton_max_iters,
                             tol=w.newton_tol)
        if mode == "run":
            st, _ = sim.solve(b.top_water, b.top_heat, nsteps, w.steps_per_slot)
            iters = st.mean_iters * nsteps
        else:
            iters = 0.0
            for s in range(nsteps):
                _set_forcing(m, b, b.forcing_slot(s))
                iters += sim.step().mean_iters
        outs.append((sim.state(), iters))
        m.close()
    (Yr, ir), (Ys, is_) = outs
    for k in Yr:
        assert np.array_equal(Yr[k], Ys[k], equal_nan=True), k
    print(f"warm-started run of {nsteps} steps equals the steps one call at a time; Newton "
          f"iterations per column {ir:.3f} / {is_:.3f}", flush=True)
    assert abs(ir - is_) <= 1e-9 * max(1.0, ir)
