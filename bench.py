#!/usr/bin/env python3
"""Benchmark of the EDUS propagation hot path on one MI355X.

A step is one RK4 step of the whole k grid (4 right-hand sides: E.grad rho by
spectral derivatives, the commutator with H0 + E.xi, the stage updates).
Workload (default): 14-band hexagonal tight-binding model, 120x120 k grid,
x-polarised pulse, dt = 0.05 a.u., FP64 state and arithmetic, FULL layout
(the reference's element-for-element arithmetic).

Prints one JSON line.  `value` = k-point * RK4 steps / s of the timed region
with the state resident in HBM.  `roofline` is for the kernel class with the
largest share of the timed region, from HIP events on the engine's stream.
`cpu_baseline` times the FP64 CPU restatement (oracle/) on this host.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "edus_2.0_amd"))

METRIC = "k-point·RK4-steps/s (nb=14, 120² k); learned-parareal vs plain RK4 @1e-6"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(cls: str, npl: int, nk: int, nb: int, n_e: int) -> float:
    """Algorithmic HBM bytes per launch: one array of npl planes is npl*nk*16 B.
    grad_col reads rho, writes G; grad_row reads rho and G, writes D; comm
    (RK4 stage average) reads rho_s, H0, n_E xi, D plus the stage's state
    arrays, writes its outputs -- (6 + n_E) arrays on average over the four
    stages; monitor reads the nb diagonal planes."""
    s = 16.0 * npl * nk
    return {"grad_col": 2 * s, "grad_row": 3 * s, "comm": (6 + n_e) * s,
            "monitor": 16.0 * nb * nk}.get(cls, 0.0)


def cpu_baseline(params: dict, steps: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = oracle_py.Oracle(params)
    o.step(1)  # warm-up
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return {"value": o.Nk * steps / dt, "unit": "k-point·RK4-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle/ FP64 restatement (3 FFTs + 2 zgemm per k per RHS), same workload and grid, "
                      f"{steps} RK4 steps after 1 warm-up step, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="ipa14_hex120")
    ap.add_argument("--layout", choices=["full", "hermitian"], default="full")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32-MFMA commutator with FP64 state, gradient and updates")
    ap.add_argument("--graph", type=int, default=20,
                    help="RK4 steps per replayed hipGraph in the timed region (0: direct launches)")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np  # noqa: F401

    import edus
    from edus import models

    tmp = tempfile.mkdtemp(prefix="edus_bench_")
    params = models.workload(args.workload, tmp)
    layout = edus.LAYOUT_FULL if args.layout == "full" else edus.LAYOUT_HERMITIAN
    prec = edus.PREC_FP64 if args.precision == "fp64" else edus.PREC_MIXED
    sim = edus.Simulation(params, layout=layout, precision=prec)
    eng = sim.engine
    nb, nk = sim.nb, sim.Nk
    npl = nb * nb if layout == edus.LAYOUT_FULL else nb * (nb + 1) // 2

    # timed region: K steps, graph replay unless --graph 0
    eng.graph(args.graph)
    sim.steps(args.warmup)
    eng.synchronize()
    eng.check()
    t0 = time.perf_counter()
    sim.steps(args.steps)
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    trace_dev, herm_dev = eng.check()

    # kernel timing: the next K steps with direct launches, every launch
    # bracketed by HIP events on the engine's stream
    eng.graph(0)
    eng.timing_enable(True)
    eng.synchronize()
    t1 = time.perf_counter()
    sim.steps(args.steps)
    eng.synchronize()
    elapsed_direct = time.perf_counter() - t1
    classes = ["grad_col", "grad_row", "comm", "hf_row", "hf_col", "monitor"]
    times = {c: eng.timing_get(c) for c in classes}
    eng.timing_enable(False)

    value = nk * args.steps / elapsed
    n_e = int(sum(1 for x in params["lasers"][0]["polarization"] if x != 0))
    dom = max(classes, key=lambda c: times[c][0])
    ms, cnt = times[dom]
    avg_s = ms / 1e3 / max(cnt, 1)
    byts = kernel_bytes(dom, npl, nk, nb, n_e)
    achieved = byts / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": byts,
                "avg_launch_us": avg_s * 1e6,
                "timed_over": f"{args.steps} direct-launch steps after the timed region (same kernels)",
                "direct_ms_per_step": elapsed_direct / args.steps * 1e3,
                "kernel_share": {c: round(times[c][0] / 1e3 / elapsed_direct, 4) for c in classes if times[c][1]},
                "avg_us": {c: round(times[c][0] * 1e3 / times[c][1], 2) for c in classes if times[c][1]},
                "bytes_per_launch_all": {c: kernel_bytes(c, npl, nk, nb, n_e) for c in classes if times[c][1]}}

    out = {
        "metric": METRIC, "value": value, "unit": "k-point·RK4-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f64 (f32 commutator)",
        "data": "synthetic",
        "config": {"workload": args.workload, "nb": nb, "grid": [sim.N0, sim.N1], "dt_au": sim.dt,
                   "integrator": "RK4", "layout": args.layout, "precision": args.precision,
                   "graph_steps": args.graph,
                   "model": "seeded random tight-binding (edus.models)", "tuning": eng.get_tuning().as_dict()},
        "roofline": roofline,
        "monitors": {"trace_dev": trace_dev, "herm_dev": herm_dev},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(params, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
