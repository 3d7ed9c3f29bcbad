"""The batched and learned-parareal workloads at their real shapes.

- 64 heterogeneous pulse runs (nb in {2, 6, 14} x 48^2 / 96^2 / 120^2, peak
  fields log-spaced over 1e-3..3e-2 a.u., random CEP and photon energy) as one
  batch with hipGraph replay: one member of each of the nine classes against
  the FP64 oracle driven by the same pulse (max|d rho| <= 1e-10), every member
  bit for bit against the same context stepped alone;
- the 8-band HSEX model on 96^2 k: GPU RK4 against the oracle at the pulse
  peak, then parareal over 32 slices of 200 steps converging to sequential
  GPU RK4 within max|d rho| <= 1e-6.
"""
import copy
import math
import os
import sys

import numpy as np
import pytest

import edus
from edus import config, models
from edus import parareal as pr

oracle_py = pytest.importorskip("oracle_py")

pytestmark = pytest.mark.gpu

HARTREE_EV = 2 * 2.1798723611035e-18 / 1.602176634e-19
AU_TIME_FS = 2.4188843265857e-2


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(p, o):
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=p["filledbands"], U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")))
    if p["coulomb"]:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    e = edus.Engine(o.nb, o.N0, o.N1, **args)
    e.set_state(o.rho)
    return e


def _as_laser(e0, omega_ev, duration_fs, phase, t0_fs):
    """The pulse as the reference's laser input (sin^2 envelope of
    duration_fs, x-polarised)."""
    las = dict(config.LASER_DEFAULTS)
    las.update(models.laser(e0, omega_ev, duration_fs, (1, 0, 0), phase))
    las["t0"], las["t0_units"] = t0_fs, "fs"
    return las


def test_hetero_batch64_at_shape(tmp_path):
    classes = [(w, g) for w in ("ipa2_square30", "ipa6_hex96", "ipa14_hex120") for g in (48, 96, 120)]
    n, steps, dt = 64, 2, 0.05
    rng = np.random.default_rng(7)
    e0 = np.logspace(np.log10(1e-3), np.log10(3e-2), n)
    omega = rng.uniform(0.8, 2.0, n)
    cep = rng.uniform(0, 2 * np.pi, n)
    dur = 20.0
    t0_fs = -dur / 2  # the steps run at the pulses' peak
    # the stage times of the two steps, from t = 0
    t = dt * np.arange(steps)[:, None] + dt * np.array([0.0, 0.5, 1.0])[None, :]
    base, oracles = {}, {}
    members, E = [], np.zeros((steps, 3, n, 3))
    for i in range(n):
        w, g = classes[i % len(classes)]
        if (w, g) not in base:
            p = models.workload(w, os.path.join(str(tmp_path), f"{w}_{g}"), grid=[g, g, 1])
            p = copy.deepcopy(p)
            # this class's first member is driven by the reference's own laser in the oracle
            p["lasers"] = [_as_laser(e0[i], omega[i], dur, cep[i], t0_fs)]
            _say("set-up", w, g)
            o = oracle_py.Oracle(p)
            base[(w, g)] = _engine(p, o)
            oracles[(w, g)] = (i, o)
        if oracles[(w, g)][0] == i:
            o = oracles[(w, g)][1]
            E[:, :, i, :] = np.array([[o.E(x) for x in row] for row in t])
        else:
            pl = pr.Pulse(e0[i], omega[i], dur, (1.0, 0.0, 0.0), cep[i], t0=t0_fs / AU_TIME_FS)
            E[:, :, i, :] = pl(t)
        members.append(base[(w, g)].clone())
    batch = edus.Batch(members)
    ng, _ = batch.groups()
    batch.graph(steps)
    batch.rk4_steps(dt, E)
    batch.synchronize()
    for i, m in enumerate(members):
        batch.store(i)
        w, g = classes[i % len(classes)]
        alone = base[(w, g)].clone()
        alone.rk4_steps(dt, E[:, :, i, :])
        assert np.array_equal(m.get_state(), alone.get_state()), (i, w, g)
        alone.close()
    for (w, g), (i, o) in oracles.items():
        o.step(steps)
        err = float(np.abs(members[i].get_state() - o.rho).max())
        moved = float(np.abs(o.rho - base[(w, g)].get_state()).max())
        _say(f"{w} {g}x{g} member {i}: max|d rho| vs oracle {err:.2e} (state moved by {moved:.2e})")
        assert err <= 1e-10 and moved > 1e-9
    _say("groups", ng)
    assert ng == 9


def test_learned_parareal_workload_at_shape(tmp_path):
    p = models.workload("hsex8_hex96", str(tmp_path))
    peak = copy.deepcopy(p)
    peak["lasers"] = [dict(las, t0=-20.0) for las in peak["lasers"]]  # 40 fs pulse: steps at its peak
    o = oracle_py.Oracle(peak)
    eng = _engine(peak, o)
    dt = o.dt
    t = dt * np.arange(4)[:, None] + dt * np.array([0.0, 0.5, 1.0])[None, :]
    E = np.array([[o.E(x) for x in row] for row in t])
    eng.rk4_steps(dt, E)
    o.step(4)
    err = float(np.abs(eng.get_state() - o.rho).max())
    _say("HSEX 8 bands 96x96, 4 RK4 steps at the pulse peak: max|d rho| vs oracle", err)
    assert err <= 1e-10

    # parareal from the ground state over the workload's own pulse
    o0 = oracle_py.Oracle(p)
    eng0 = _engine(p, o0)
    N, FINE, COARSE, DT = 32, 200, 10, 0.05
    field = lambda tt: np.array([o0.E(x) for x in np.asarray(tt).ravel()]).reshape(np.shape(tt) + (3,))
    Ef, Ec, Es = pr.slice_tables(field, N, FINE, COARSE, DT)
    seq = eng0.clone()
    bounds = [seq.get_state()]
    for k in range(N):
        seq.rk4_steps(DT, Ef[k])
        bounds.append(seq.get_state())
    para = pr.Parareal(eng0, N, FINE, COARSE, DT, tol=1e-6, max_iter=6, graph_steps=50)
    para.set_fields(Ef, Ec, Es)
    res = para.run()
    derr = max(float(np.abs(para.state(k) - bounds[k]).max()) for k in range(N + 1))
    _say("parareal 32 x 200 steps:", res, "max|d rho| vs sequential", derr)
    assert res["converged"] == 1 and res["fallback"] == 0
    # the coarse sweep starts from the slice boundaries just written (a stale
    # boundary buffer makes the first update O(1))
    assert len(res["updates"]) == res["iterations"] and res["updates"][0] < 0.1
    assert derr <= 1e-6
