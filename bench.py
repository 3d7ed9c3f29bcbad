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


def kernel_bytes(cls: str, npl: int, nk: int, nb: int, n_e: int, solver: str = "RK", order: int = 4,
                 coulomb: bool = False, peierls: bool = False, f32_fields: bool = False,
                 tri_fields: bool = False, hf_fused: bool = False) -> float:
    """Algorithmic HBM bytes per launch: one array of npl planes is
    S = npl*nk*16 B.  grad_col reads rho, writes G (2S); grad_row reads rho
    and G, writes D (3S); grad_fused (both axes in one pass per plane) reads
    rho, writes D (2S); kgrad (finite-difference stencil, gradient_space
    "k") reads rho, writes D (2S); hf_row (forward or inverse row transform) 2S;
    hf_col (column convolution with the real W) 2.5S.  comm reads rho_s, H0,
    n_E xi, D (+ the exchange self energy) and the stage's state arrays and
    writes its outputs: RK4's stages move (2, 2, 4, 2) S beyond the 3 + n_E
    operand arrays, (5.5 + n_E) S on average (the second stage leaves the
    accumulator alone; the third recovers its term from rho - y), AB of
    order p reads y and p-1 history arrays and writes y and the new history
    entry, (4 + n_E + p) S.  Peierls gauge: peierls_col reads H0 and the
    n_E xi and writes the half-phased
    planes ((2 + n_E) S), peierls_row reads those and writes S_A[H] (2S), and
    comm reads S_A[H] instead of H0, xi and D (one array in place of 2 + n_E).
    f32_fields: the mixed-precision commutator reads H0 and the n_E xi from
    FP32 copies, half an array each.  tri_fields (FULL layout): it reads
    them from their upper triangles, nb (nb + 1) / 2 of the nb^2 planes
    (0.536 of an array each at nb = 14).  hf_fused (exchange): grad_row also
    writes the exchange term's forward row transform (4S), and hf_row then
    counts only the inverse row passes (2S each)."""
    s = 16.0 * npl * nk
    sig = 1 if coulomb else 0
    comm = (5.5 + n_e + sig) if solver == "RK" else (4 + n_e + order + sig)
    if peierls:
        comm -= 1 + n_e
    field = 1.0  # bytes of H0 and of each xi, in arrays
    if f32_fields:
        field *= 0.5
    if tri_fields:
        field *= (nb + 1) / (2.0 * nb)
    if not peierls:
        comm -= (1.0 - field) * (1 + n_e)
    return {"grad_col": 2 * s, "grad_row": (4 if (coulomb and hf_fused) else 3) * s, "grad_fused": 2 * s, "kgrad": 2 * s, "comm": comm * s,
            "hf_row": 2 * s,
            "hf_col": 2.5 * s, "peierls_col": (2 + n_e) * s,
            "peierls_row": 2 * s}.get(cls, 0.0)


def cpu_baseline(params: dict, steps: int, integ: str, bytes_per_kstep: float = 0.0, budget_s: float = 18.0):
    """The FP64 CPU restatement (oracle/) on this host's cores, same workload
    and grid: warm-up steps, then three timed runs, the median of which is
    reported (steps <= 0: each run as many steps as fill budget_s / 3).  The
    host's copy bandwidth is measured beside it, with the rate the byte model
    (bytes_per_kstep per k-point and step) allows at that bandwidth."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import statistics

    import oracle_py

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = oracle_py.Oracle(params)
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    warm = int(max(1, min(20, 3.0 / max(one, 1e-6))))
    o.step(warm - 1)
    if steps <= 0:
        steps = int(max(3, min(2000, budget_s / 3 / max(one, 1e-6))))
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.step(steps)
        runs.append(time.perf_counter() - t0)
    med = statistics.median(runs)
    bw = oracle_py.host_bandwidth(1 << 28, 3)
    out = {"value": o.Nk * steps / med, "unit": f"k-point·{integ}-steps/s", "cores": threads, "kind": "port",
           "sample": f"oracle/ FP64 restatement (3 FFTs + 2 zgemm per k per RHS), same workload and grid: "
                     f"{warm} warm-up steps, then the median of 3 runs of {steps} {integ} steps "
                     f"({', '.join(f'{r:.2f}' for r in runs)} s)",
           "runs_s": runs, "host_copy_GBs": bw / 1e9}
    if bytes_per_kstep > 0:
        out["byte_model_bound"] = bw / bytes_per_kstep
    return out


def measured_traffic(cls: str, workload: str, layout: str, precision: str, gradient: str = "R",
                     gauge: str = "length", fields: str = "full"):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    counter summaries under profiles/ (FETCH_SIZE and WRITE_SIZE runs scaled
    by the copy-kernel calibration); None when no summary covers the case.
    A summary without a gradient=, gauge= or fields= key covers the defaults
    (the spectral gradient, the length gauge, H0 and xi read from every
    element) only."""
    import glob

    want = {"class": cls, "workload": workload, "layout": layout, "precision": precision,
            "gradient": gradient, "gauge": gauge, "fields": fields}
    default = {"gradient": "R", "gauge": "length", "fields": "full"}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.txt"))):
        with open(path) as f:
            for line in f:
                if not line.startswith("traffic:"):
                    continue
                kv = dict(x.split("=", 1) for x in line.split()[1:])
                if all(kv.get(k, default.get(k)) == v for k, v in want.items()):
                    return float(kv["bytes_per_launch"]), os.path.relpath(path, ROOT)
    return None, None


def graph_chunk(steps: int, graph: int) -> int:
    """Steps per replayed graph: the largest divisor of the timed step count
    that is at most `graph` (0: direct launches)."""
    if graph <= 0 or steps <= 0:
        return 0
    for g in range(min(graph, steps), 0, -1):
        if steps % g == 0:
            return g
    return 1


def warmup_steps(warmup: int, gsteps: int) -> int:
    """Untimed steps before t0: at least `warmup`, rounded up to whole graphs
    and at least one graph, so the timed steps only replay."""
    if gsteps <= 0:
        return warmup
    return max(gsteps, -(-warmup // gsteps) * gsteps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="ipa14_hex120")
    ap.add_argument("--layout", choices=["full", "hermitian"], default="full")
    ap.add_argument("--peierls", action="store_true",
                    help="Peierls gauge (the reference's peierls: true); steps use direct launches")
    ap.add_argument("--gradient-space", choices=["R", "k"], default="R",
                    help="R: spectral derivative (the reference default); k: finite-difference stencil")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32-MFMA commutator with FP64 state, gradient and updates")
    ap.add_argument("--graph", type=int, default=20,
                    help="RK4 steps per replayed hipGraph in the timed region (0: direct launches)")
    ap.add_argument("--cpu-steps", type=int, default=0, help="CPU baseline steps (0: about 15 s worth)")
    ap.add_argument("--coarse-steps", type=int, default=20, help="parareal: coarse RK4 steps per slice")
    ap.add_argument("--train-runs", type=int, default=64, help="parareal: random-pulse training runs")
    ap.add_argument("--k-samples", type=int, default=256, help="parareal: k-points sampled per run and slice")
    ap.add_argument("--epochs", type=int, default=20, help="parareal: training epochs")
    ap.add_argument("--tol", type=float, default=1e-6, help="parareal: convergence tolerance")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg of the default line")
    ap.add_argument("--no-parareal", action="store_true", help="skip the learned-parareal leg of the default line")
    ap.add_argument("--no-batch", action="store_true", help="skip the heterogeneous-batch leg of the default line")
    ap.add_argument("--batch-steps", type=int, default=40, help="timed steps of each mode of the batch leg")
    ap.add_argument("--parareal-train-runs", type=int, default=16,
                    help="random-pulse training runs of the parareal leg in the default line")
    ap.add_argument("--tuning", default=None,
                    help="JSON of edus_tuning fields to force after the autotuner (e.g. for counter runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the k-grid is not sharded across GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "hetero_batch64":
        run_batch(args)
        return
    if args.workload == "parareal_hsex8_96":
        run_parareal(args)
        return
    if args.workload == "screened_ipa14_hex120":
        run_screened(args)
        return
    if args.workload == "rotate_ipa14_hex120":
        run_rotate(args)
        return
    if args.workload == "hbm_copy":
        run_copy(args)
        return

    import numpy as np  # noqa: F401

    import edus
    from edus import models

    tmp = tempfile.mkdtemp(prefix="edus_bench_")
    params = models.workload(args.workload, tmp)
    params["gradient_space"] = args.gradient_space
    params["peierls"] = bool(args.peierls)
    layout = edus.LAYOUT_FULL if args.layout == "full" else edus.LAYOUT_HERMITIAN
    prec = edus.PREC_FP64 if args.precision == "fp64" else edus.PREC_MIXED
    sim = edus.Simulation(params, layout=layout, precision=prec)
    eng = sim.engine
    nb, nk = sim.nb, sim.Nk
    npl = nb * nb if layout == edus.LAYOUT_FULL else nb * (nb + 1) // 2
    if args.tuning:
        tun = eng.get_tuning()
        for key, val in json.loads(args.tuning).items():
            if isinstance(val, list):
                for i, v in enumerate(val):
                    getattr(tun, key)[i] = int(v)
            else:
                setattr(tun, key, int(val))
        eng.set_tuning(tun)

    # timed region: K steps, graph replay unless --graph 0.  The graph holds
    # the largest divisor of K up to --graph steps, and the warm-up runs whole
    # graphs (at least one), so capture and instantiation stay outside t0..t1.
    gsteps = graph_chunk(args.steps, args.graph)
    warm = warmup_steps(args.warmup, gsteps)
    eng.graph(gsteps)
    sim.steps(warm)
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
    classes = ["grad_col", "grad_row", "grad_fused", "kgrad", "comm", "hf_row", "hf_col", "peierls_col",
               "peierls_row"]
    times = {c: eng.timing_get(c) for c in classes}
    eng.timing_enable(False)

    value = nk * args.steps / elapsed
    integ = "RK4" if params["solver"] == "RK" else f"AB{params['order']}"
    pol = np.zeros(3)
    for las in params["lasers"]:
        pol += np.abs(np.asarray(las["polarization"], dtype=float))
    n_e = int(np.count_nonzero(pol))
    solver, order, coul = params["solver"], int(params["order"]), bool(params["coulomb"])
    f32 = args.precision == "mixed" and not coul and not args.peierls  # the engine's FP32 copies of H0 and xi
    tun = eng.get_tuning()
    tri = bool(tun.tri_fields)  # H0 and xi read from their upper triangles
    hf = coul and bool(tun.hf_fused) and tun.fft_family[1] == 1  # exchange row pass from the gradient's
    kb = {c: kernel_bytes(c, npl, nk, nb, n_e, solver, order, coul, bool(args.peierls), f32, tri, hf)
          for c in classes}
    dom = max(classes, key=lambda c: times[c][0])
    ms, cnt = times[dom]
    avg_s = ms / 1e3 / max(cnt, 1)
    byts = kb[dom]
    achieved = byts / avg_s / 1e9
    traffic, traffic_src = measured_traffic(dom, args.workload, args.layout, args.precision, args.gradient_space,
                                            "peierls" if args.peierls else "length", "tri" if tri else "full")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                "bytes_per_launch": byts,
                "avg_launch_us": avg_s * 1e6,
                "timed_over": f"{args.steps} direct-launch steps after the timed region (same kernels)",
                "direct_ms_per_step": elapsed_direct / args.steps * 1e3,
                "kernel_share": {c: round(times[c][0] / 1e3 / elapsed_direct, 4) for c in classes if times[c][1]},
                "avg_us": {c: round(times[c][0] * 1e3 / times[c][1], 2) for c in classes if times[c][1]},
                "bytes_per_launch_all": {c: kb[c] for c in classes if times[c][1]}}

    out = {
        "metric": METRIC, "value": value, "unit": f"k-point·{integ}-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f64 (f32 commutator)",
        "data": "synthetic",
        "config": {"workload": args.workload, "nb": nb, "grid": [sim.N0, sim.N1], "dt_au": sim.dt,
                   "integrator": integ, "layout": args.layout, "precision": args.precision,
                   "gradient_space": args.gradient_space, "peierls": bool(args.peierls),
                   "graph_steps": 0 if args.peierls else gsteps,
                   "warmup_steps_run": warm,
                   "model": "seeded random tight-binding (edus.models)", "tuning": eng.get_tuning().as_dict()},
        "roofline": roofline,
        "monitors": {"trace_dev": trace_dev, "herm_dev": herm_dev},
    }
    if args.precision == "fp64" and not args.peierls and not args.no_mixed:
        out["mixed"] = mixed_leg(params, layout, args)
        _progress("mixed precision", out["mixed"])
    if args.workload == "ipa14_hex120" and not args.no_batch:
        out["batch"] = batch_leg(args.batch_steps, 20, 20)
        _progress("batch", {k: out["batch"][k] for k in ("value", "batch_over_sequential",
                                                          "batch_over_best_concurrent")})
    if args.workload == "ipa14_hex120" and not args.no_parareal:
        out["parareal"] = parareal_leg(args.parareal_train_runs, args.k_samples, args.epochs, args.tol,
                                       args.coarse_steps)
    if not args.no_cpu_baseline:
        bps = 16.0 * nb * nb * (36 + 4 * n_e) if (solver == "RK" and not coul) else 0.0
        out["cpu_baseline"] = cpu_baseline(params, args.cpu_steps, integ, bps)
    print(json.dumps(out), flush=True)


def mixed_leg(params: dict, layout: int, args) -> dict:
    """The same workload in mixed precision (FP32-MFMA commutator; FP64
    gradient, state and stage updates) against FP64, both started at the
    pulse's peak (laser t0 moved back by half the pulse) from the ground
    state: the mixed run's throughput (graph replay, as the headline) and,
    after the same warm-up + K steps, max|d rho| and the relative current
    error against the FP64 engine (which the tests pin to the oracle)."""
    import copy

    import numpy as np

    import edus

    p = copy.deepcopy(params)
    for las in p["lasers"]:
        period_fs = 4.135667696 / float(las["frequency"])  # h / (hbar omega in eV), fs
        las["t0"] = -0.5 * float(las["cycles"]) * period_fs
        las["t0_units"] = "fs"
    res = {}
    gsteps = graph_chunk(args.steps, args.graph)
    warm = warmup_steps(args.warmup, gsteps)
    for prec in (edus.PREC_MIXED, edus.PREC_FP64):
        sim = edus.Simulation(p, layout=layout, precision=prec)
        eng = sim.engine
        eng.graph(gsteps)
        sim.steps(warm)
        eng.synchronize()
        t0 = time.perf_counter()
        sim.steps(args.steps)
        eng.synchronize()
        el = time.perf_counter() - t0
        _, _, vel = eng.observables()
        E = sim.field(sim.time)
        res[prec] = (eng.get_state(), np.asarray(vel), el, sim.Nk, E)
        sim.close()
    rm, vm, el, nk, E = res[edus.PREC_MIXED]
    rf, vf, elf, _, _ = res[edus.PREC_FP64]
    return {"value": nk * args.steps / el, "ms_per_step": el / args.steps * 1e3,
            "fp64_value_same_leg": nk * args.steps / elf, "steps_compared": warm + args.steps,
            "field_at_end_au": [float(x) for x in E],
            "max_drho_vs_fp64": float(np.abs(rm - rf).max()),
            "rel_current_err_vs_fp64": float(np.abs(vm - vf).max() / max(np.abs(vf).max(), 1e-300)),
            "note": "commutator on FP32 MFMA with FP32 accumulation; gradient, state and stage updates FP64"}


def _progress(*a):
    """One line per finished phase, to stderr."""
    print("bench:", *a, file=sys.stderr, flush=True)


def run_batch(args):
    """The c3 workload alone, as its own line (see batch_leg)."""
    print(json.dumps(batch_leg(args.steps, args.warmup, args.graph)), flush=True)


def batch_leg(steps: int, warmup: int, graph: int) -> dict:
    """Heterogeneous batch: 64 pulse runs over nb in {2, 6, 14} and grids
    48^2, 96^2, 120^2 (one seeded model per class), peak fields log-spaced
    over 1e-3 .. 3e-2 a.u., CEP uniform in [0, 2 pi), photon energy uniform in
    [0.8, 2.0] eV.  The same number of steps runs four ways: sequentially
    (one member after another, each with graph replay of graph_steps steps),
    as 64 independent contexts launched back to back (each on its own stream,
    so they overlap) with graph replay and with direct launches, and as one
    batch (stacked groups, one hipGraph per graph_steps steps); throughput =
    sum over members of Nk * steps / s.  The direct-launch contexts and the
    batch are timed in alternating rounds and their medians compared (each
    advances its own copy of the states)."""
    import statistics

    import numpy as np

    import edus
    from edus import models
    from edus.parareal import Pulse, stage_table

    rng = np.random.default_rng(7)
    classes = [(w, g) for w in ("ipa2_square30", "ipa6_hex96", "ipa14_hex120") for g in (48, 96, 120)]
    tmp = tempfile.mkdtemp(prefix="edus_batch_")
    sims = {}
    for w, g in classes:
        p = models.workload(w, os.path.join(tmp, f"{w}_{g}"), grid=[g, g, 1])
        sims[(w, g)] = edus.Simulation(p)
        _progress("set up", w, g)
    n = 64
    e0 = np.logspace(np.log10(1e-3), np.log10(3e-2), n)
    members, pulses, nks = [], [], []
    for i in range(n):
        w, g = classes[i % len(classes)]
        members.append(sims[(w, g)].engine.clone())
        pulses.append(Pulse(e0[i], rng.uniform(0.8, 2.0), 20.0, (1.0, 0.0, 0.0), rng.uniform(0, 2 * np.pi)))
        nks.append(g * g)
    dt, K, W = 0.05, steps, max(warmup, graph)
    rounds = 3
    t_start = 200.0  # inside the pulses
    E = np.stack([stage_table(pl, t_start, dt, 4 * W + (2 + 2 * rounds) * K) for pl in pulses],
                 axis=2)  # [s][j][i][3]
    off = 0

    def chunk(c):  # the next c steps of every member's field table
        nonlocal off
        e = E[off:off + c]
        off += c
        return e

    def contexts(g, c):  # every member on its own stream, launched back to back (they overlap)
        for m in members:
            m.graph(g)
        e = chunk(c)
        t0 = time.perf_counter()
        for i, m in enumerate(members):
            m.rk4_steps(dt, e[:, :, i, :])
        for m in members:
            m.synchronize()
        return time.perf_counter() - t0

    # sequential: one member at a time, graph replay
    for m in members:
        m.graph(graph)
    e = chunk(W)
    for i, m in enumerate(members):
        m.rk4_steps(dt, e[:, :, i, :])
        m.synchronize()
    e = chunk(K)
    t0 = time.perf_counter()
    for i, m in enumerate(members):
        m.rk4_steps(dt, e[:, :, i, :])
        m.synchronize()
    t_seq = time.perf_counter() - t0
    _progress("sequential", t_seq, "s")
    contexts(graph, W)
    t_conc = contexts(graph, K)
    _progress("concurrent contexts, graph replay", t_conc, "s")
    # concurrent contexts with direct launches and the batch, alternated
    # (rounds of each, medians reported: a single short run of either varies by ~10%)
    contexts(0, W)
    batch = edus.Batch(members)
    ng, _ = batch.groups()
    batch.graph(graph)
    batch.rk4_steps(dt, chunk(W))
    batch.synchronize()
    t_cd, t_b = [], []
    for _ in range(rounds):  # the batch and the contexts advance their own copies of the states
        t_cd.append(contexts(0, K))
        e = chunk(K)
        t1 = time.perf_counter()
        batch.rk4_steps(dt, e)
        batch.synchronize()
        t_b.append(time.perf_counter() - t1)
    t_conc_d = statistics.median(t_cd)
    t_batch = statistics.median(t_b)
    _progress("concurrent contexts, direct launches", t_cd, "s; batch", t_b, "s")
    tr, herm = batch.check()
    work = float(sum(nks)) * K
    out = {
        "metric": "k-point·RK4-steps/s, 64 heterogeneous pulse runs", "value": work / t_batch,
        "unit": "k-point·RK4-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": t_batch / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "hetero_batch64", "members": n, "groups": ng, "graph_steps": graph,
                   "classes": [f"{w} {g}x{g}" for w, g in classes]},
        "rounds": rounds,
        "rounds_ms_per_step": {"batch": [t / K * 1e3 for t in t_b],
                               "concurrent_contexts_direct": [t / K * 1e3 for t in t_cd]},
        "sequential": {"value": work / t_seq, "ms_per_step": t_seq / K * 1e3},
        "concurrent_contexts_graph": {"value": work / t_conc, "ms_per_step": t_conc / K * 1e3},
        "concurrent_contexts_direct": {"value": work / t_conc_d, "ms_per_step": t_conc_d / K * 1e3},
        "batch_over_sequential": t_seq / t_batch,
        "batch_over_best_concurrent": min(t_conc, t_conc_d) / t_batch,
        "monitors": {"trace_dev": tr, "herm_dev": herm},
    }
    batch.close()
    for m in members:
        m.close()
    return out



def run_copy(args):
    """Counter calibration: a plain elementwise copy (b = 1.0 * a, one read
    and one write of 256 MiB per launch) through torch, so that rocprofv3's
    FETCH_SIZE / WRITE_SIZE per dispatch can be scaled to bytes moved."""
    import torch

    n = (256 << 20) // 8
    a = torch.arange(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(args.warmup, 1)):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        torch.mul(a, 1.0, out=b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    nbytes = 2.0 * n * 8
    print(json.dumps({"metric": "plain copy (counter calibration)", "value": nbytes / (ms * 1e-3) / 1e9,
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
                      "higher_is_better": True, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "hbm_copy", "read_bytes": n * 8, "write_bytes": n * 8}}), flush=True)


def run_rotate(args):
    """The Wannier -> Bloch rotation U^dagger rho U of every k-point (the
    reference's go_to_bloch, src/Operator/Operator.hpp:528-564) at the 14-band
    120^2 shape: K rotations of the device state into the k-major staging
    buffer (edus_get_state_bloch_device: the kernel behind
    edus_get_state_bloch, without the host copy), launched back to back and
    each timed by HIP events on the engine's stream.  Algorithmic bytes per
    launch: the operator and U read, the result written, 3 x nb^2 x Nk x 16 B
    (HBM-bound: 2 complex nb x nb products per k are ~4.7 flop/B, under the
    FP64 ridge)."""
    import numpy as np

    import edus
    from edus import models

    tmp = tempfile.mkdtemp(prefix="edus_rot_")
    params = models.workload("ipa14_hex120", tmp)
    prec = edus.PREC_FP64 if args.precision == "fp64" else edus.PREC_MIXED
    sim = edus.Simulation(params, precision=prec)
    eng = sim.engine
    nb, nk = sim.nb, sim.Nk
    from edus import hip_rt  # the HIP runtime the engine links (not torch's copy)

    out = hip_rt.DeviceBuffer(nk * nb * nb * 16)  # [Nk][nb][nb] complex
    for _ in range(max(args.warmup, 1)):
        eng.get_state_bloch_device(out.address)
    eng.synchronize()
    eng.timing_enable(True)
    for _ in range(args.steps):  # back to back on the context's stream
        eng.get_state_bloch_device(out.address)
    eng.synchronize()
    ms, cnt = eng.timing_get("rotate")
    eng.timing_enable(False)
    # the device result against the host path's (the same kernel)
    ref = eng.get_state_bloch()
    got = out.to_array(np.complex128, ref.shape)
    if not np.array_equal(got, ref):
        raise RuntimeError("edus_get_state_bloch_device differs from edus_get_state_bloch")
    avg_s = ms / 1e3 / max(cnt, 1)
    byts = 3.0 * nb * nb * nk * 16
    achieved = byts / avg_s / 1e9
    traffic, traffic_src = measured_traffic("rotate", "rotate_ipa14_hex120", "full", args.precision)
    print(json.dumps({"metric": "k-point gauge rotations/s (U^dagger rho U, nb=14, 120^2 k)",
                      "value": nk / avg_s, "unit": "k-point rotations/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": avg_s * 1e3, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None,
                      "dtype": "f64" if args.precision == "fp64" else "f32 (FP32 MFMA, FP32 accumulation)",
                      "data": "synthetic",
                      "config": {"workload": "rotate_ipa14_hex120", "nb": nb, "grid": [sim.N0, sim.N1],
                                 "precision": args.precision, "launches_timed": cnt},
                      "roofline": {"bound": "hbm", "kernel": "rotate", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                                   "traffic_source": traffic_src, "bytes_per_launch": byts,
                                   "avg_launch_us": avg_s * 1e6,
                                   "timed_over": f"HIP events around each of {cnt} back-to-back launches "
                                                 "(edus_get_state_bloch_device) on the context's stream"}}),
          flush=True)


def lindhard_flops(nk: int, nb: int, nf: int) -> float:
    """Algorithmic FP64 flops of one Lindhard response: for every pair (q, k),
    element a <= b (X_ba = conj X_ab) and conduction band c: B = conj(u_a) u_b
    (6 flops), S = sum_v r_vc A_v (4 nv), acc += S B (8)."""
    nc = nb - nf
    return float(nk) ** 2 * (nb * (nb + 1) // 2) * nc * (14 + 4 * nf)


def run_screened(args):
    """The screened-potential application's Lindhard response
    (Screened_potential.cpp:17-58) on the 14-band hexagonal model's 120^2 grid:
    one step = one full response X(q) for all q (Nk^2 (q, k) pairs), inputs
    (eigenvectors, energies) resident in HBM, timed with HIP events."""
    import numpy as np

    import edus
    from edus import hip_rt, models

    tmp = tempfile.mkdtemp(prefix="edus_bench_")
    params = models.workload("ipa14_hex120", tmp)
    sim = edus.Simulation(params)
    nb, n0, n1, nk, nf = sim.nb, sim.N0, sim.N1, sim.Nk, int(params["filledbands"])
    U, eps = sim.get("U"), sim.get("eps")
    sim.close()
    dU = hip_rt.DeviceBuffer.from_array(U)
    de = hip_rt.DeviceBuffer.from_array(np.ascontiguousarray(eps))
    work = hip_rt.DeviceBuffer(U.nbytes)
    dX = hip_rt.DeviceBuffer(U.nbytes)
    call = lambda: edus.lindhard_device(nb, n0, n1, nf, dU.address, de.address, work.address, dX.address, 0)
    for _ in range(max(args.warmup, 1)):
        call()
    hip_rt.synchronize()
    t = hip_rt.Timer()
    t.start()
    for _ in range(args.steps):
        call()
    ms = t.stop_ms() / args.steps
    flops = lindhard_flops(nk, nb, nf)
    out = {
        "metric": "Lindhard response (q, k) pairs/s (screened potential, nb=14, 120² k)",
        "value": nk * nk / (ms * 1e-3), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic 14-band model (seed 1), eigenvectors and energies from the host set-up",
        "config": {"workload": "screened_ipa14_hex120", "nb": nb, "grid": [n0, n1], "filledbands": nf},
        "roofline": {"bound": "fp64", "achieved": flops / (ms * 1e-3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": flops / (ms * 1e-3) / 1e12 / 78.6, "traffic": None,
                     "kernel": "lindhard_kernel", "note": "includes the band-split pass (under 1% of the call)"},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        g = 32  # same model on a 32 x 32 grid: the reference's loop has the same cost per (q, k) pair
        small = edus.Simulation(models.workload("ipa14_hex120", tmp, grid=[g, g, 1]))
        Us, es = small.get("U"), small.get("eps")
        small.close()
        oracle_py.lindhard(Us[:4], es[:4], 2, 2, nf)  # load
        t0 = time.perf_counter()
        oracle_py.lindhard(Us, es, g, g, nf)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": (g * g) ** 2 / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
                               "sample": f"oracle/ restatement of the reference's loop (q, k, a, b, c, v) on the same "
                                         f"model at {g}x{g} k, {(g * g) ** 2} pairs in {dt:.2f} s"}
    print(json.dumps(out), flush=True)


def parareal_leg(train_runs: int, k_samples: int, epochs: int, tol: float, coarse: int) -> dict:
    """Learned parareal on the 8-band HSEX model, 96^2 k: 32 slices of 200
    fine RK4 steps (dt 0.05 a.u.) of the pulse of the hsex8_hex96 workload.
    Training data: `train_runs` runs with random pulses of the plain method;
    the network is trained on the GPU.  Reported: wall time of sequential GPU
    RK4 (graph replay), of parareal with the coarse RK4 alone (the method
    without the network) and with the network, iterations, and max|d rho| of
    both against sequential RK4 at every slice boundary."""
    import numpy as np

    import edus
    from edus import models
    from edus import parareal as pr

    N, FINE, COARSE, DT = 32, 200, coarse, 0.05
    tmp = tempfile.mkdtemp(prefix="edus_para_")
    params = models.workload("hsex8_hex96", tmp)
    sim = edus.Simulation(params)
    eng = sim.engine
    _progress("parareal: set up")

    def field(t):
        t = np.asarray(t, dtype=float)
        return np.array([sim.field(x) for x in t.ravel()]).reshape(t.shape + (3,))

    Ef, Ec, Es = pr.slice_tables(field, N, FINE, COARSE, DT)
    # sequential GPU RK4
    seq = eng.clone()
    seq.graph(50)
    seq.rk4_steps(DT, Ef[0][:50])  # capture outside the timed region
    seq.set_state(eng.get_state())
    seq.synchronize()
    t0 = time.perf_counter()
    for n in range(N):
        seq.rk4_steps(DT, Ef[n])
    seq.synchronize()
    t_seq = time.perf_counter() - t0
    _progress("parareal: sequential RK4", t_seq, "s")
    # slice boundaries of the sequential run (untimed)
    chk = eng.clone()
    bounds = [chk.get_state()]
    for n in range(N):
        chk.rk4_steps(DT, Ef[n])
        bounds.append(chk.get_state())
    chk.close()
    seq.close()
    # training data and training (one-off costs, reported separately)
    t1 = time.perf_counter()
    pulses = pr.random_pulses(train_runs, seed=11)
    tabs = [pr.slice_tables(pl, N, FINE, COARSE, DT) for pl in pulses]
    P = pr.Params(N, FINE, COARSE, DT, tol, 6, 50)
    X, T = pr.collect(eng, P, np.stack([t[0] for t in tabs]), np.stack([t[1] for t in tabs]),
                      np.stack([t[2] for t in tabs]), k_samples=k_samples, seed=5)
    t_collect = time.perf_counter() - t1
    _progress("parareal: training data", X.shape, t_collect, "s")
    t2 = time.perf_counter()
    stats = pr.normalisation(X, T)
    net, hist = pr.train_network(X, T, stats, hidden=256, epochs=epochs, batch=1024, lr=1e-3, seed=7)
    t_train = time.perf_counter() - t2
    _progress("parareal: training", t_train, "s, loss", float(hist[0]), "->", float(hist[-1]))

    def run(with_net):
        para = pr.Parareal(eng, N, FINE, COARSE, DT, tol=tol, max_iter=6, graph_steps=50)
        para.set_fields(Ef, Ec, Es)
        if with_net:
            para.set_network(net, stats)
        para.run()  # warm-up: graphs captured, tiles tuned
        res = para.run()
        err = max(float(np.abs(para.state(n) - bounds[n]).max()) for n in range(N + 1))
        res["max_drho_vs_sequential"] = err
        para.close()
        _progress("parareal", "with" if with_net else "without", "network", res)
        return res

    plain = run(False)
    learned = run(True)
    return {"workload": "parareal_hsex8_96", "slices": N, "fine_steps": FINE, "coarse_steps": COARSE, "dt_au": DT,
            "tol": tol, "max_iter": 6, "train_runs": train_runs, "k_samples": k_samples, "epochs": epochs,
            "network": f"{X.shape[1]}-256-256-{T.shape[1]}",
            "sequential_rk4_s": t_seq, "parareal_plain": plain, "parareal_learned": learned,
            "learned_over_sequential": learned["seconds"] / t_seq,
            "learned_over_plain_iterations": [learned["iterations"], plain["iterations"]],
            "training": {"rows": int(X.shape[0]), "collect_s": t_collect, "train_s": t_train,
                         "loss_first": float(hist[0]), "loss_last": float(hist[-1])}}


def run_parareal(args):
    """The learned-parareal leg alone, as its own line (time-like metric)."""
    r = parareal_leg(args.train_runs, args.k_samples, args.epochs, args.tol, args.coarse_steps)
    learned = r["parareal_learned"]
    out = {
        "metric": "learned-parareal wall time vs plain GPU RK4 (8-band HSEX, 96^2 k, 32 x 200 steps)",
        "value": learned["seconds"], "unit": "s", "n_gpus": 1, "steps": r["slices"] * r["fine_steps"], "warmup": 0,
        "ms_per_step": learned["seconds"] / (r["slices"] * r["fine_steps"]) * 1e3, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64 (network bf16/f32)", "data": "synthetic",
        "config": {k: r[k] for k in ("workload", "slices", "fine_steps", "coarse_steps", "dt_au", "tol", "max_iter",
                                     "train_runs", "k_samples", "epochs", "network")},
    }
    out.update({k: r[k] for k in ("sequential_rk4_s", "parareal_plain", "parareal_learned", "training")})
    out["speedup_learned_vs_sequential"] = r["sequential_rk4_s"] / learned["seconds"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
