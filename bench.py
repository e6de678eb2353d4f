"""Benchmark: column-steps/s of the implicit soil-column step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--form F]

A "step" is one ARS111 implicit step (explicit stage + Newton to the
workload's tolerance) over every column of the workload.  Inputs are
synthetic (clmsoil.inputs) and resident in HBM before timing; the K steps
run as one captured HIP graph.  Prints one JSON line.

Forms (EnergyHydrology): "fp64" = the plain FP64 path with the reference's
modified-Picard Jacobian; "mixed" = that Jacobian and its solve in FP32 inside
an FP64-residual Newton; "newton" = the mixed form with the exact Jacobian of
the same residual (CLM_JACOBIAN_NEWTON, FP32); "learned" = the learned Newton
warm start on top of "newton".  RichardsModel: "fp64" and "learned" (on the
FP64 path).  `value` is the chosen form's throughput over all column-steps;
`converged_value` counts only the column-steps that met the Newton tolerance
and stayed finite.  The learned form's network is trained first (untimed) on
columns drawn from other seeds of the same workload: plain steps of all its
columns, spread evenly over the timed horizon, give the (state, converged
step) pairs (2,097,152 of them).  The line also carries the 131,072-column
RichardsModel workload (FP64 and learned) under "rre_131072_uneven".
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "climaland.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

# HBM bytes per step of the step kernels measured with rocprofv3 (FETCH_SIZE
# and WRITE_SIZE summed over the bins' launches, FETCH scaled by the factor a
# plain device copy of known size reads for it, WRITE likewise); raw sums
# beside them.  Figures, date, build and commands:
# profiles/k_eh_coop_eh_262144_uneven.md.
MEASURED_TRAFFIC = {("eh_262144_uneven", "learned"): {"scaled": 2836.4e6, "raw": 2008.4e6,
                                                      "source": "rocprofv3 2026-12-03, build "
                                                                "b795bf6, 3 early steps"},
                    ("eh_262144_uneven", "newton"): {"scaled": 2507.1e6, "raw": 1843.9e6,
                                                     "source": "rocprofv3 2026-12-03, build "
                                                               "b795bf6, 3 early steps"}}

# What bounds the fused step: VALU issue.  rocprofv3 counters of the step
# kernels on 3 early steps of eh_262144_uneven, learned form
# (profiles/k_eh_coop_eh_262144_uneven.md): SQ_INSTS_VALU_FLOPS_FP64 x 64
# lanes per step, SQ_INSTS_VALU per step, and the time per step of the same
# workload's bin launches in a kernel trace (first start to last end: the
# bins of a step run concurrently).
MEASURED_VALU = {"eh_262144_uneven": {"fp64_flop_per_step": 1.9933e10,
                                      "valu_instructions_per_step": 4.685e8,
                                      "kernel_ms_per_step": 1.2798,
                                      "source": "rocprofv3 2026-12-03, build b795bf6, 3 early "
                                                "steps, learned form"}}
FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X FP64 vector
CU_COUNT, ENGINE_GHZ = 256, 2.4


def step_bytes(batch, learned: bool) -> float:
    """Compulsory HBM bytes of one fused step over all columns (the group
    kernels read the column state, faces, parameters and forcing once and
    write the state once; everything else stays in registers).  Per column:
    EnergyHydrology 56N + 176, RichardsModel 24N + 96; the learned form adds
    the per-cell guess the step kernel reads (16N EH, 8N RRE)."""
    N = batch.nlayers.astype(np.float64)
    if batch.model == 1:
        per = 56 * N + 176 + (16 * N if learned else 0)
    else:
        per = 24 * N + 96 + (8 * N if learned else 0)
    return float(per.sum())


def mlp_flops(batch) -> float:
    """Multiply-add FLOPs of one warm-start inference over all columns."""
    F, Fp, H, No = (105, 128, 384, 64) if batch.model == 1 else (40, 64, 256, 32)
    return 2.0 * batch.ncol * (Fp * H + 2 * H * H + H * No)


def cpu_baseline(batch, nsteps: int, ncol_sample: int, threads: int):
    from oracle.oracle import OracleModel
    w = batch.workload
    sub = batch.column_subset(np.arange(ncol_sample))
    o = OracleModel.from_batch(sub)
    Y = sub.copy_state()
    t0 = time.perf_counter()
    o.run(Y, sub.top_water, sub.top_heat, w.steps_per_slot, w.dt, nsteps, w.newton_max_iters,
          w.newton_tol, nthreads=threads, want_iters=False)
    el = time.perf_counter() - t0
    return ncol_sample * nsteps / el, el


def parse_shapes(spec):
    """"16:16x1,24:32x1" -> [(16, 16, 1), (24, 32, 1)] (bucket max layers: G x R)."""
    out = []
    for item in filter(None, (spec or "").split(",")):
        mx, gr = item.split(":")
        g, r = gr.split("x")
        out.append((int(mx), int(g), int(r)))
    return out


def train_warm_start(name, ncol, jac, epochs, horizon, collect=8, batch=8192, jacobian=0):
    """Collect (state, converged step) pairs with the plain solver on columns
    of other seeds -- `collect` steps spread evenly over the run's `horizon`
    -- and train the warm-start network on the device."""
    import torch
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    bt = inputs.make_batch(inputs.training_workload(name), ncol=ncol, nsteps=horizon)
    w = bt.workload
    m = model_from_batch(bt, jac_precision=jac, solver=1, jacobian=jacobian)
    ws = NewtonWarmStart(m, seed=5)
    sim = LandSimulation(m, bt.copy_state(), w.dt, max_iters=w.newton_max_iters,
                         tol=w.newton_tol)
    t0 = time.perf_counter()
    stride = max(1, horizon // collect)
    rows = ws.collect(sim, bt.top_water, bt.top_heat, stride * collect, w.steps_per_slot,
                      stride)
    t1 = time.perf_counter()
    losses = ws.train(epochs, batch=batch, lr=1e-3, seed=9)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    params = ws.params()
    m.close()
    info = {"pairs": rows, "collected_every": stride, "epochs": epochs, "batch": batch,
            "params": int(params.size),
            "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
            "collect_s": t1 - t0, "train_s": t2 - t1}
    print(f"[bench] warm start: {rows} pairs collected in {t1 - t0:.1f} s, {epochs} epochs in "
          f"{t2 - t1:.1f} s, loss {losses[0]:.4g} -> {losses[-1]:.4g}", file=sys.stderr,
          flush=True)
    return params, info


def run_form(batch, jac, steps, warmup, solver=0, shapes_spec=None, ws_params=None,
             jacobian=0):
    import torch
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    w = batch.workload
    m = model_from_batch(batch, jac_precision=jac, solver=solver, jacobian=jacobian)
    if ws_params is not None:
        ws = NewtonWarmStart(m)
        ws.set_params(ws_params)
        ws.enable(True)
    sim = LandSimulation(m, batch.copy_state(), w.dt, max_iters=w.newton_max_iters,
                         tol=w.newton_tol)
    sim.set_forcing_table(batch.top_water, batch.top_heat)
    shapes = None
    if m.shapes() and shapes_spec:  # group solver, shapes given
        for mx, g, r in parse_shapes(shapes_spec):
            m.set_shape(mx, g, r)
        shapes = [s for s in m.shapes() if s[3] > 0]
    elif m.shapes():  # group solver: autotune the per-depth shapes on the initial state
        m.boundary_conditions["top"] = _top_bc(batch)
        m.set_forcing()
        sim._push()
        m.autotune(w.dt, reps=2)
        shapes = [s for s in m.shapes() if s[3] > 0]
    if shapes:
        print(f"[bench] shapes (max layers, lanes/column, rows/lane, columns): "
              f"{shapes}", file=sys.stderr, flush=True)
    if warmup > 0:  # untimed warmup steps
        sim.solve(None, None, warmup, w.steps_per_slot, upload=False)
    # capture the K-step graph outside the timed region, then restart from u0
    sim.reset(batch.copy_state())
    sim.solve(None, None, steps, w.steps_per_slot, upload=False)
    sim.reset(batch.copy_state())
    sim._push()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, _ = sim.solve(None, None, steps, w.steps_per_slot, upload=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the same (deterministic) run again to copy out its iteration table
    sim.reset(batch.copy_state())
    _, iters = sim.solve(None, None, steps, w.steps_per_slot, want_iters=True, upload=False)
    graph_ms = st.ms / steps
    infer_ms = 0.0
    if ws_params is not None:  # device time of the inference part of each step
        sim.reset(batch.copy_state())
        sim._push()
        m.boundary_conditions["top"] = _top_bc(batch)
        m.set_forcing()
        infer_ms = ws.inference_ms(w.dt, reps=5)
    m.close()
    return dict(wall=wall, stats=st, iters=iters, graph_ms=graph_ms, infer_ms=infer_ms,
                kernel_ms=graph_ms - infer_ms, shapes=shapes)


def _top_bc(batch):
    from clmsoil.model import HeatFluxBC, WaterFluxBC, WaterHeatBC
    if batch.top_heat is None:
        return WaterFluxBC(batch.top_water[0])
    return WaterHeatBC(WaterFluxBC(batch.top_water[0]), HeatFluxBC(batch.top_heat[0]))


# (jac_precision, jacobian) of each form; "learned" adds the warm start to
# the form named in LEARNED_ON
FORMS = {"fp64": (64, 0), "mixed": (32, 0), "newton": (32, 1)}
LEARNED_ON = {1: "newton", 0: "fp64"}  # by model: EnergyHydrology, RichardsModel


def form_summary(batch, steps, r):
    st = r["stats"]
    n = batch.ncol * steps
    return {"value": n / r["wall"], "ms_per_step": r["wall"] * 1e3 / steps,
            "converged_value": (n - st.n_nonconverged) / r["wall"],
            "mean_newton_iters": st.mean_iters,
            "nonconverged_fraction": st.n_nonconverged / n,
            "nonfinite_fraction": st.n_nan / n,
            "kernel_ms_per_step": r["kernel_ms"],
            "warm_start_inference_ms_per_step": r["infer_ms"] or None}


def run_forms(batch, steps, warmup, names, args, epochs):
    """Run the named forms (untimed training first for "learned"); returns
    ({form: run_form result}, training info)."""
    w = batch.workload
    res, ws_info = {}, None
    for name in names:
        if name == "learned":
            jac, jacobian = FORMS[LEARNED_ON[batch.model]]
            collect = 16 if batch.model == 0 else 8  # 2,097,152 pairs either way
            params, ws_info = train_warm_start(w.name, batch.ncol, jac, epochs, steps,
                                               collect=collect, jacobian=jacobian)
        else:
            jac, jacobian = FORMS[name]
            params = None
        r = run_form(batch, jac, steps, warmup, args.solver, args.shapes, params, jacobian)
        res[name] = r
        st = r["stats"]
        print(f"[bench] {w.name} {name}: {batch.ncol * steps / r['wall']:.4g} column-steps/s, "
              f"{r['wall'] * 1e3 / steps:.3f} ms/step, mean Newton iters {st.mean_iters:.2f}, "
              f"non-converged {st.n_nonconverged}, non-finite {st.n_nan}",
              file=sys.stderr, flush=True)
    return res, ws_info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="eh_262144_uneven")
    ap.add_argument("--form", default="learned", choices=["learned", "newton", "mixed", "fp64"])
    ap.add_argument("--epochs", type=int, default=16, help="warm-start training epochs")
    ap.add_argument("--ncol", type=int, default=None, help="prefix of the workload (testing)")
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rre", action="store_true",
                    help="leave out the rre_131072_uneven key")
    ap.add_argument("--solver", type=int, default=0, help="0 auto, 1 group, 2 lane")
    ap.add_argument("--shapes", default=None,
                    help='fix group shapes instead of autotuning, e.g. "16:16x1,24:32x1"')
    ap.add_argument("--only", action="store_true",
                    help="run the chosen form only (no other forms)")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per step measured with rocprofv3 (FETCH_SIZE, WRITE_SIZE, "
                         "scaled by a copy kernel) to report beside the compulsory bytes")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the timed steps, run a plain device copy of MIB MiB 5 times "
                         "(reference kernel for HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; columns shard across GPUs with no "
              "exchange, so N>1 is not run here", file=sys.stderr)
        sys.exit(2)
    from clmsoil import inputs
    batch = inputs.make_batch(args.workload, ncol=args.ncol)
    w = batch.workload
    steps = args.steps or w.nsteps
    batch_steps = inputs.make_batch(args.workload, ncol=args.ncol, nsteps=steps) \
        if steps != w.nsteps else batch
    print(f"[bench] workload {w.name}: {batch.ncol} columns, {batch.ncell} cells, {steps} "
          f"steps", file=sys.stderr, flush=True)
    if batch.model == 1:
        names = ["fp64", "mixed", "newton", "learned"]
    else:
        names = ["fp64", "learned"]
    if args.form not in names:
        sys.exit(f"bench.py: form {args.form} is not defined for {w.name}")
    if args.only:
        names = [args.form]
    res, ws_info = run_forms(batch_steps, steps, args.warmup, names, args, args.epochs)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: 5 copies of {n * 8} bytes", file=sys.stderr,
              flush=True)
    r = res[args.form]
    st = r["stats"]
    value = batch.ncol * steps / r["wall"]
    learned = args.form == "learned"
    jac = FORMS[LEARNED_ON[batch.model] if learned else args.form][0]
    bytes_step = step_bytes(batch_steps, learned)
    achieved = bytes_step / (r["kernel_ms"] * 1e-3) / 1e9
    kname = "k_eh_coop" if batch.model == 1 else "k_rre_coop"
    ncs = batch.ncol * steps
    traffic = MEASURED_TRAFFIC.get((w.name, args.form), {})
    out = {
        "metric": "column-steps/s at Newton residual 1e-8: learned+mixed vs plain FP64 GPU",
        "value": value,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if jac == 64 else "f64 residual/state, f32 jacobian+solve") +
                 (", bf16 warm-start network (fp32 accumulate)" if learned else ""),
        "data": "synthetic",
        "config": {"workload": w.name, "columns": batch.ncol, "layers": list(w.layers),
                   "dt_s": w.dt, "newton_tol": w.newton_tol,
                   "newton_max_iters": w.newton_max_iters,
                   "form": {"learned": f"learned warm start + {LEARNED_ON[batch.model]} form",
                            "newton": "mixed precision, exact Newton matrix",
                            "mixed": "mixed precision (FP32 Jacobian + solve)",
                            "fp64": "plain FP64"}[args.form],
                   "solver": "group" if r.get("shapes") else "lane",
                   "shapes": r.get("shapes")},
        "converged_value": (ncs - st.n_nonconverged) / r["wall"],
        "plain_fp64_value": (ncs / res["fp64"]["wall"]) if "fp64" in res else None,
        "speedup_vs_plain_fp64": (value / (ncs / res["fp64"]["wall"]))
        if "fp64" in res else None,
        "other_forms": {k: form_summary(batch_steps, steps, v) for k, v in res.items()},
        "warm_start_training": ws_info,
        "warm_start_inference": None if not learned else {
            "ms_per_step": r["infer_ms"],
            "gemm_tflops": mlp_flops(batch_steps) / (r["infer_ms"] * 1e-3) / 1e12,
            "mfma_bf16_peak_tflops": 2500.0,
            "note": "clm_ws_time: k_features then k_ws_infer, HIP events on the context's "
                    "stream; gemm_tflops = the four layers' FLOPs over both kernels' time "
                    "(k_ws_infer alone: profiles/k_ws_infer_eh_262144_uneven.md)"},
        "newton": {"mean_iters": st.mean_iters, "max_iters": st.max_iters_used,
                   "nonconverged_column_steps": st.n_nonconverged,
                   "nonfinite_column_steps": st.n_nan,
                   "nonconverged_fraction": st.n_nonconverged / ncs,
                   "nonfinite_fraction": st.n_nan / ncs,
                   "note": "non-converged includes non-finite; a non-finite column stops "
                           "iterating in the step it turns non-finite and in later steps"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": args.traffic if args.traffic is not None else
                     traffic.get("scaled"),
                     "traffic_raw": traffic.get("raw"),
                     "traffic_source": traffic.get("source"),
                     "kernel": kname + " (one launch per depth bin or bin pair; per step)",
                     "kernel_ms_per_step": r["kernel_ms"],
                     "bytes_per_step": bytes_step,
                     "note": "fused step: achieved = compulsory HBM bytes / kernel time; the "
                             "kernel is bound by FP64 VALU issue and latency, not HBM"},
    }
    v = MEASURED_VALU.get(w.name)
    if v is not None:
        t = v["kernel_ms_per_step"] * 1e-3
        out["valu_roofline"] = {
            "kernel": kname, "source": v["source"],
            "fp64_tflops": v["fp64_flop_per_step"] / t / 1e12,
            "fp64_peak_tflops": FP64_VALU_PEAK_TFLOPS,
            "fp64_frac": v["fp64_flop_per_step"] / t / 1e12 / FP64_VALU_PEAK_TFLOPS,
            "valu_issue_per_cu_cycle": v["valu_instructions_per_step"] /
            (CU_COUNT * t * ENGINE_GHZ * 1e9),
            "valu_issue_peak_per_cu_cycle": 1.0}
    if batch.model == 1 and not args.no_rre and args.ncol is None:
        rb = inputs.make_batch("rre_131072_uneven")
        rsteps = rb.workload.nsteps
        rres, rinfo = run_forms(rb, rsteps, args.warmup, ["fp64", "learned"], args, args.epochs)
        rbytes = step_bytes(rb, False)
        out["rre_131072_uneven"] = {
            "steps": rsteps,
            "forms": {k: form_summary(rb, rsteps, v) for k, v in rres.items()},
            "roofline_fp64": {"bound": "valu", "kernel": "k_rre_coop",
                              "achieved": rbytes / (rres["fp64"]["kernel_ms"] * 1e-3) / 1e9,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": rbytes / (rres["fp64"]["kernel_ms"] * 1e-3) / 1e9
                              / HBM_PEAK_GBS, "bytes_per_step": rbytes},
            "warm_start_training": rinfo}
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        # about 12 s of CPU work at the oracle's ~5e5 column-steps/s on 16 threads
        default = max(256, min(batch.ncol, (int(5.9e6 / steps) // 256) * 256))
        sample = args.cpu_sample or default
        cps, el = cpu_baseline(batch_steps, steps, sample, threads)
        # one core on a sixteenth of the sample, beside the reference's own
        # single-process CPU figures
        sample1 = max(256, sample // 16)
        cps1, el1 = cpu_baseline(batch_steps, steps, sample1, 1)
        out["cpu_baseline"] = {"value": cps, "unit": "column-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"first {sample} columns x {steps} steps of {w.name}, "
                                         f"FP64 oracle (reference Jacobian), OpenMP over "
                                         f"columns ({el:.1f} s)",
                               "single_core": {"value": cps1, "cores": 1,
                                               "sample": f"first {sample1} columns x {steps} "
                                                         f"steps, one thread ({el1:.1f} s)"}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
