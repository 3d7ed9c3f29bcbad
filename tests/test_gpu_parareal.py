"""GPU: learned parareal (edus_parareal_*) on a small 8-band HSEX model.

- plain parareal (coarse RK4 alone) converges to sequential RK4 within the
  max|d rho| <= 1e-6 bar;
- with max_iter = 0 the fallback is sequential RK4 itself, bit for bit;
- a network whose output scale is zero gives the plain run bit for bit;
- the per-k network, trained on the GPU on runs with random pulses, removes
  most of the coarse propagator's error on held-out data, and learned parareal
  converges to sequential RK4 within the bar; against the ablation (coarse RK4
  alone) its first sweep is closer to the fine solution and it needs no more
  iterations (at this small shape; the 96^2 c4 measurement is in the bench).
"""
import os
import sys

import numpy as np
import pytest

import edus
from edus import models
from edus import parareal as pr

pytestmark = pytest.mark.gpu

N, FINE, COARSE, DT = 8, 40, 4, 0.05


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("parareal"))
    p = models.workload("hsex8_hex96", os.path.join(d, "m"), grid=[12, 12, 1])
    sim = edus.Simulation(p)
    pulse = pr.Pulse(0.01, 1.8, 10.0, (1.0, 0.0, 0.0), 0.3)
    Ef, Ec, Es = pr.slice_tables(pulse, N, FINE, COARSE, DT)
    # sequential RK4 reference with every slice boundary
    ref = sim.engine.clone()
    rho0 = ref.get_state()
    bounds = [rho0]
    for n in range(N):
        ref.rk4_steps(DT, Ef[n])
        bounds.append(ref.get_state())
    return sim, (Ef, Ec, Es), bounds


def _err(a, b):
    return float(np.abs(a - b).max())


def test_plain_parareal_converges_to_sequential(setup):
    sim, (Ef, Ec, Es), bounds = setup
    para = pr.Parareal(sim.engine, N, FINE, COARSE, DT, tol=1e-9, max_iter=N)
    para.set_fields(Ef, Ec, Es)
    res = para.run()
    errs = [_err(para.state(n), bounds[n]) for n in range(N + 1)]
    _say("plain parareal", res, "max|d rho| per boundary", max(errs))
    assert res["converged"] == 1 and res["fallback"] == 0
    assert res["iterations"] <= N
    assert max(errs) <= 1e-6
    # the same run from an explicit initial state (uploaded), and a second run
    # of the same object, end bit for bit where the first did
    states = [para.state(n) for n in range(N + 1)]
    for rho0 in (bounds[0], None):
        again = para.run(rho0=rho0)
        assert again["iterations"] == res["iterations"] and again["updates"] == res["updates"]
        for n in range(N + 1):
            assert np.array_equal(para.state(n), states[n]), n


def test_zero_scaled_network_is_plain_parareal(setup):
    """The network's correction enters as y * tsig + tmu: with both zero the
    learned run is the plain one bit for bit (same updates, same boundaries),
    whatever the untrained network computes -- the features, the network and
    the correction's unpacking run, and add exact zeros."""
    sim, (Ef, Ec, Es), bounds = setup
    eng = sim.engine
    q = eng.nb * eng.nb
    din = 2 * q + 2 * Es.shape[1] + 4
    net = edus.Mlp(din, 64, q, seed=2)
    stats = {"mu": np.zeros(din, np.float32), "sig": np.ones(din, np.float32),
             "tmu": np.zeros(q, np.float32), "tsig": np.zeros(q, np.float32)}
    runs = []
    for with_net in (False, True):
        para = pr.Parareal(eng, N, FINE, COARSE, DT, tol=1e-9, max_iter=N)
        para.set_fields(Ef, Ec, Es)
        if with_net:
            para.set_network(net, stats)
        res = para.run()
        runs.append((res, [para.state(n) for n in range(N + 1)]))
        para.close()
    (r0, s0), (r1, s1) = runs
    assert r0["iterations"] == r1["iterations"] and r0["updates"] == r1["updates"]
    for a, b in zip(s0, s1):
        assert np.array_equal(a, b)


def test_fallback_is_sequential_rk4(setup):
    sim, (Ef, Ec, Es), bounds = setup
    para = pr.Parareal(sim.engine, N, FINE, COARSE, DT, tol=1e-12, max_iter=0)
    para.set_fields(Ef, Ec, Es)
    res = para.run()
    assert res["fallback"] == 1 and res["converged"] == 0
    for n in range(N + 1):
        assert np.array_equal(para.state(n), bounds[n]), n


def test_learned_parareal_and_ablation(setup):
    sim, (Ef, Ec, Es), bounds = setup
    eng = sim.engine
    # training runs: random pulses, same slicing
    pulses = pr.random_pulses(12, seed=3, duration_fs=(5.0, 15.0))
    tabs = [pr.slice_tables(pl, N, FINE, COARSE, DT) for pl in pulses]
    params = pr.Params(N, FINE, COARSE, DT, 1e-9, N, 0)
    X, T = pr.collect(eng, params, np.stack([t[0] for t in tabs]), np.stack([t[1] for t in tabs]),
                      np.stack([t[2] for t in tabs]), k_samples=eng.Nk, seed=1)
    assert np.isfinite(X).all() and np.isfinite(T).all()
    rows = X.shape[0]
    cut = rows * 10 // 12  # the last two runs are held out
    stats = pr.normalisation(X[:cut], T[:cut])
    net, hist = pr.train_network(X[:cut], T[:cut], stats, hidden=128, epochs=60, batch=512, lr=2e-3, seed=4)
    Yv = net.forward((X[cut:] - stats["mu"]) / stats["sig"]) * stats["tsig"] + stats["tmu"]
    base = float(np.mean(T[cut:].astype(np.float64) ** 2))      # coarse RK4 alone
    left = float(np.mean((T[cut:] - Yv).astype(np.float64) ** 2))  # coarse RK4 + network
    _say("training loss first/last", hist[0], hist[-1], "held-out coarse error^2 without / with network", base, left)
    assert hist[-1] < hist[0]
    assert left < 0.5 * base

    plain = pr.Parareal(eng, N, FINE, COARSE, DT, tol=1e-9, max_iter=N)
    plain.set_fields(Ef, Ec, Es)
    r0 = plain.run()
    learned = pr.Parareal(eng, N, FINE, COARSE, DT, tol=1e-9, max_iter=N)
    learned.set_fields(Ef, Ec, Es)
    learned.set_network(net, stats)
    r1 = learned.run()
    err = max(_err(learned.state(n), bounds[n]) for n in range(N + 1))
    _say("ablation: iterations without network", r0["iterations"], "with network", r1["iterations"],
         "learned max|d rho|", err, "results", r0, r1)
    assert r1["converged"] == 1 and err <= 1e-6
    # here the network's first sweep is closer to the fine solution than
    # coarse RK4's, and the learned run needs no more iterations
    assert r1["updates"][0] < r0["updates"][0]
    assert r1["iterations"] <= r0["iterations"]
