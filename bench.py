"""Benchmark: column-steps/s of the implicit soil-column step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--form F]

A "step" is one ARS111 implicit step (explicit stage + Newton to the
workload's tolerance) over every column of the workload.  Inputs are
synthetic (clmsoil.inputs) and resident in HBM before timing; the K steps
run as one captured HIP graph.  Prints one JSON line.

Forms: "learned" = the learned Newton warm start on top of "mixed";
"mixed" = FP32 Jacobian + FP32 solve inside an FP64-residual Newton;
"fp64" = the plain FP64 path.  `value` is the chosen form's throughput; the
plain FP64 throughput on the same workload is reported beside it.  The
learned form's network is trained first (untimed) on columns drawn from other
seeds of the same workload: 8 plain steps of all its columns, spread evenly
over the timed horizon, give the (state, converged step) pairs.
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
# and WRITE_SIZE summed over the bins' launches, FETCH scaled x2.0 and WRITE
# x1.0 by a plain 1 GiB device copy); raw sums beside them.  They include the
# scratch writes of the register spills (outside the Newton loop) of the
# mixed-form kernels, which are compiled for four waves per SIMD.  Summary and
# commands: profiles/k_eh_coop_eh_262144_uneven.md.
MEASURED_TRAFFIC = {("eh_262144_uneven", "learned"): {"scaled": 1718.7e6, "raw": 1116.7e6},
                    ("eh_262144_uneven", "mixed"): {"scaled": 1469.1e6, "raw": 985.3e6}}

# What bounds the fused step: VALU issue.  rocprofv3 counters of the step
# kernels on 3 early steps of eh_262144_uneven, mixed form, shapes
# (16,1),(32,1),(32,1),(64,1) (profiles/k_eh_coop_eh_262144_uneven.md):
# SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes per step, SQ_INSTS_VALU per step, and
# the kernels' time per step in the same build's trace.
MEASURED_VALU = {"eh_262144_uneven": {"fp64_flop_per_step": 4.604e10,
                                      "valu_instructions_per_step": 1.046e9,
                                      "kernel_ms_per_step": 2.496}}
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


def unfused_bytes(batch, iters: np.ndarray) -> float:
    """Bytes the unfused hook-by-hook decomposition would move for the same
    Newton iteration counts ([steps][ncol]): explicit stage plus, per
    iteration, assemble (aux + residual + Jacobian) and solve, each streaming
    its operands through HBM (EH 80N+136 + k(280N+96); RRE k(104N+72)+64)."""
    N = batch.nlayers.astype(np.float64)[None, :]
    k = (iters & 0xFFFF).astype(np.float64)
    if batch.model == 1:
        per = 80 * N + 136 + k * (280 * N + 96)
    else:
        per = k * (104 * N + 72) + 64
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


def train_warm_start(name, ncol, jac, epochs, horizon, collect=8, batch=8192):
    """Collect (state, converged step) pairs with the plain solver on columns
    of other seeds -- `collect` steps spread evenly over the run's `horizon`
    -- and train the warm-start network on the device."""
    import torch
    from clmsoil import inputs
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    bt = inputs.make_batch(inputs.training_workload(name), ncol=ncol, nsteps=horizon)
    w = bt.workload
    m = model_from_batch(bt, jac_precision=jac, solver=1)
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


def run_form(batch, jac, steps, warmup, solver=0, shapes_spec=None, ws_params=None):
    import torch
    from clmsoil.model import LandSimulation, NewtonWarmStart, model_from_batch
    w = batch.workload
    m = model_from_batch(batch, jac_precision=jac, solver=solver)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="eh_262144_uneven")
    ap.add_argument("--form", default="learned", choices=["learned", "mixed", "fp64"])
    ap.add_argument("--epochs", type=int, default=16, help="warm-start training epochs")
    ap.add_argument("--ncol", type=int, default=None, help="prefix of the workload (testing)")
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solver", type=int, default=0, help="0 auto, 1 group, 2 lane")
    ap.add_argument("--shapes", default=None,
                    help='fix group shapes instead of autotuning, e.g. "16:16x1,24:32x1"')
    ap.add_argument("--only", action="store_true",
                    help="run the chosen form only (no plain FP64 comparison)")
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
    forms = {"learned": 32, "mixed": 32, "fp64": 64}
    if batch.model == 0:  # RichardsModel: the learned form runs on the FP64 Jacobian
        forms["learned"] = 64
    chosen = forms[args.form]
    res = {}
    ws_params, ws_info = None, None
    if args.form == "learned":
        ws_params, ws_info = train_warm_start(w.name, batch.ncol, chosen, args.epochs, steps)
    for name, jac in (("fp64", 64), ("mixed", 32), ("learned", forms["learned"])):
        if name != args.form and (args.only or (batch.model == 0 and name == "mixed")):
            continue
        if name == "learned" and ws_params is None:
            continue
        r = run_form(batch_steps, jac, steps, args.warmup, args.solver, args.shapes,
                     ws_params if name == "learned" else None)
        res[name] = r
        st = r["stats"]
        print(f"[bench] {name}: {batch.ncol * steps / r['wall']:.4g} column-steps/s, "
              f"{r['wall'] * 1e3 / steps:.3f} ms/step, mean Newton iters {st.mean_iters:.2f}, "
              f"non-converged {st.n_nonconverged}, non-finite {st.n_nan}",
              file=sys.stderr, flush=True)
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
    bytes_step = step_bytes(batch_steps, learned)
    achieved = bytes_step / (r["kernel_ms"] * 1e-3) / 1e9
    unfused = unfused_bytes(batch_steps, r["iters"]) / steps
    kname = "k_eh_coop" if batch.model == 1 else "k_rre_coop"
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
        "dtype": ("f64" if chosen == 64 else "f64 residual/state, f32 jacobian+solve") +
                 (", bf16 warm-start network (fp32 accumulate)" if args.form == "learned" else ""),
        "data": "synthetic",
        "config": {"workload": w.name, "columns": batch.ncol, "layers": list(w.layers),
                   "dt_s": w.dt, "newton_tol": w.newton_tol,
                   "newton_max_iters": w.newton_max_iters,
                   "form": {"learned": "learned warm start + mixed precision",
                            "mixed": "mixed precision (FP32 Jacobian + solve)",
                            "fp64": "plain FP64"}[args.form],
                   "solver": "group" if r.get("shapes") else "lane",
                   "shapes": r.get("shapes")},
        "plain_fp64_value": (batch.ncol * steps / res["fp64"]["wall"]) if "fp64" in res
        else None,
        "speedup_vs_plain_fp64": (value / (batch.ncol * steps / res["fp64"]["wall"]))
        if "fp64" in res else None,
        "other_forms": {k: {"value": batch.ncol * steps / v["wall"],
                            "ms_per_step": v["wall"] * 1e3 / steps,
                            "mean_newton_iters": v["stats"].mean_iters}
                        for k, v in res.items()},
        "warm_start_training": ws_info,
        "warm_start_inference": None if not learned else {
            "ms_per_step": r["infer_ms"],
            "gemm_tflops": mlp_flops(batch_steps) / (r["infer_ms"] * 1e-3) / 1e12,
            "mfma_bf16_peak_tflops": 2500.0},
        "newton": {"mean_iters": st.mean_iters, "max_iters": st.max_iters_used,
                   "nonconverged_column_steps": st.n_nonconverged,
                   "nonfinite_column_steps": st.n_nan,
                   "converged_fraction": 1.0 - (st.n_nonconverged / (batch.ncol * steps))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": args.traffic if args.traffic is not None else
                     MEASURED_TRAFFIC.get((w.name, args.form), {}).get("scaled"),
                     "traffic_raw": MEASURED_TRAFFIC.get((w.name, args.form), {}).get("raw"),
                     "kernel": kname + " (one launch per depth bin; per step)",
                     "kernel_ms_per_step": r["kernel_ms"],
                     "bytes_per_step": bytes_step,
                     "note": "fused step: compulsory bytes only; the kernel is bound by FP64 "
                             "VALU issue, not HBM",
                     "unfused_equivalent_GBps": unfused / (r["kernel_ms"] * 1e-3) / 1e9,
                     "unfused_equivalent_frac": unfused / (r["kernel_ms"] * 1e-3) / 1e9
                     / HBM_PEAK_GBS},
    }
    v = MEASURED_VALU.get(w.name)
    if v is not None:
        t = v["kernel_ms_per_step"] * 1e-3
        out["valu_roofline"] = {
            "kernel": kname, "source": "rocprofv3 counters, 3 early steps, mixed form",
            "fp64_tflops": v["fp64_flop_per_step"] / t / 1e12,
            "fp64_peak_tflops": FP64_VALU_PEAK_TFLOPS,
            "fp64_frac": v["fp64_flop_per_step"] / t / 1e12 / FP64_VALU_PEAK_TFLOPS,
            "valu_issue_per_cu_cycle": v["valu_instructions_per_step"] /
            (CU_COUNT * t * ENGINE_GHZ * 1e9),
            "valu_issue_peak_per_cu_cycle": 1.0}
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sample = args.cpu_sample or min(batch.ncol, 24576 if batch.model == 1 else 16384)
        cps, el = cpu_baseline(batch_steps, steps, sample, threads)
        # one core on a sixteenth of the sample, beside the reference's own
        # single-process CPU figures
        sample1 = max(256, sample // 16)
        cps1, el1 = cpu_baseline(batch_steps, steps, sample1, 1)
        out["cpu_baseline"] = {"value": cps, "unit": "column-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"first {sample} columns x {steps} steps of {w.name}, "
                                         f"FP64 oracle, OpenMP over columns ({el:.1f} s)",
                               "single_core": {"value": cps1, "cores": 1,
                                               "sample": f"first {sample1} columns x {steps} "
                                                         f"steps, one thread ({el1:.1f} s)"}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
