"""Benchmark: column-steps/s of the implicit soil-column step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--form F]

A "step" is one ARS111 implicit step (explicit stage + Newton to the
workload's tolerance) over every column of the workload.  Inputs are
synthetic (clmsoil.inputs) and resident in HBM before timing; the K steps
run as one captured HIP graph.  Prints one JSON line.

Forms: "mixed" = FP32 Jacobian + FP32 solve inside an FP64-residual Newton;
"fp64" = the plain FP64 path.  `value` is the chosen form's throughput; the
plain FP64 throughput on the same workload is reported beside it.
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


def design_bytes(batch, iters: np.ndarray, jac_precision: int) -> float:
    """HBM bytes the fused step kernel moves for the given per-column-step
    Newton iteration counts ([steps][ncol]).  Per-level figures follow the
    loads/stores of k_rre_step / k_eh_step."""
    j = 4 if jac_precision == 32 else 8
    N = batch.nlayers.astype(np.float64)[None, :]
    k = (iters & 0xFFFF).astype(np.float64)
    if batch.model == 0:
        per_col = N * 8 + k * N * (40 + 4 * j) + 80.0
    else:
        per_col = N * (32 + 72 + j) + k * N * (208 + 13 * j) + 152.0
    return float(per_col.sum())


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


def run_form(batch, jac, steps, warmup, solver=0, shapes_spec=None):
    import torch
    from clmsoil.model import LandSimulation, model_from_batch
    w = batch.workload
    m = model_from_batch(batch, jac_precision=jac, solver=solver)
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
    kernel_ms = st.ms / steps
    m.close()
    return dict(wall=wall, stats=st, iters=iters, kernel_ms=kernel_ms, shapes=shapes)


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
    ap.add_argument("--form", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--ncol", type=int, default=None, help="prefix of the workload (testing)")
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solver", type=int, default=0, help="0 auto, 1 group, 2 lane")
    ap.add_argument("--shapes", default=None,
                    help='fix group shapes instead of autotuning, e.g. "16:16x1,24:32x1"')
    ap.add_argument("--only", action="store_true",
                    help="run the chosen form only (no plain FP64 comparison)")
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
    forms = {"mixed": 32, "fp64": 64}
    chosen = forms[args.form]
    res = {}
    for name, jac in (("fp64", 64), ("mixed", 32)):
        if name != args.form and (args.only or (batch.model == 0 and args.form == "fp64")):
            continue
        r = run_form(batch_steps, jac, steps, args.warmup, args.solver, args.shapes)
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
    bytes_total = design_bytes(batch_steps, r["iters"], chosen)
    achieved = bytes_total / steps / (r["kernel_ms"] * 1e-3) / 1e9
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
        "dtype": "f64" if chosen == 64 else "f64 residual/state, f32 jacobian+solve",
        "data": "synthetic",
        "config": {"workload": w.name, "columns": batch.ncol, "layers": list(w.layers),
                   "dt_s": w.dt, "newton_tol": w.newton_tol,
                   "newton_max_iters": w.newton_max_iters,
                   "form": args.form + (" (no learned warm start yet)" if args.form == "mixed"
                                        else ""),
                   "solver": "group" if r.get("shapes") else "lane",
                   "shapes": r.get("shapes")},
        "plain_fp64_value": (batch.ncol * steps / res["fp64"]["wall"]) if "fp64" in res
        else None,
        "speedup_vs_plain_fp64": (value / (batch.ncol * steps / res["fp64"]["wall"]))
        if "fp64" in res else None,
        "newton": {"mean_iters": st.mean_iters, "max_iters": st.max_iters_used,
                   "nonconverged_column_steps": st.n_nonconverged,
                   "nonfinite_column_steps": st.n_nan,
                   "converged_fraction": 1.0 - (st.n_nonconverged / (batch.ncol * steps))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_eh_step" if batch.model == 1 else "k_rre_step",
                     "kernel_ms_per_launch": r["kernel_ms"],
                     "bytes_per_launch": bytes_total / steps},
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sample = args.cpu_sample or min(batch.ncol, 2048 if batch.model == 1 else 4096)
        cps, el = cpu_baseline(batch_steps, steps, sample, threads)
        out["cpu_baseline"] = {"value": cps, "unit": "column-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"first {sample} columns x {steps} steps of {w.name}, "
                                         f"FP64 oracle, OpenMP over columns ({el:.1f} s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
