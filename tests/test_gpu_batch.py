"""GPU: batches of contexts (edus_batch_*).

A batch stacks members of equal shape into groups and advances all of them
per launch; members of different shapes run as separate groups.  Every member
must end exactly where the same context ends when stepped alone with
edus_rk4_steps (the same kernels and arithmetic, field tables instead of
kernel arguments), with direct launches and with hipGraph replays.  The single
context path is the one pinned against the oracle in test_gpu_parity.
"""
import os
import sys

import numpy as np
import pytest

import edus
from edus import models

oracle_py = pytest.importorskip("oracle_py")

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# (workload, grid, seed): two groups of two members with different models,
# one member with exchange, one with 14 bands
MEMBERS = [("ipa2_square30", [12, 15], 11), ("ipa2_square30", [12, 15], 12), ("hsex8_hex90", [9, 11], 13),
           ("hsex8_hex90", [9, 11], 14), ("ipa14_hex120", [10, 9], 15)]


def _setups(model_dir):
    out = []
    for name, grid, seed in MEMBERS:
        p = models.workload(name, os.path.join(model_dir, f"b_{name}_{grid[0]}x{grid[1]}_{seed}"),
                            grid=[grid[0], grid[1], 1], seed=seed)
        o = oracle_py.Oracle(p)
        out.append((p, o))
    return out


def _engine(p, o):
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=p["filledbands"], U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")))
    if p["coulomb"]:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    e = edus.Engine(o.nb, o.N0, o.N1, **args)
    e.set_state(o.rho)
    return e


def _fields(nsteps, n, dt, seed):
    """Smooth, member-specific fields at t, t+dt/2, t+dt: E[s][j][i][3]."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(1e-3, 8e-3, size=(n, 3)) * np.array([1.0, 1.0, 0.0])
    w = rng.uniform(0.02, 0.08, size=n)
    ph = rng.uniform(0, 2 * np.pi, size=n)
    E = np.zeros((nsteps, 3, n, 3))
    for s in range(nsteps):
        for j, frac in enumerate((0.0, 0.5, 1.0)):
            t = (s + frac) * dt
            E[s, j] = amp * np.sin(w * t + ph)[:, None]
    return E


@pytest.mark.parametrize("graph_steps", [0, 4])
def test_batch_matches_single_contexts(model_dir, graph_steps):
    setups = _setups(model_dir)
    n, dt, nsteps = len(setups), 0.05, 10
    E = _fields(nsteps, n, dt, 5)
    # each member alone
    alone = []
    for i, (p, o) in enumerate(setups):
        e = _engine(p, o)
        e.rk4_steps(dt, E[:, :, i, :])
        alone.append(e.get_state())
        e.close()
    # all members in one batch
    members = [_engine(p, o) for p, o in setups]
    b = edus.Batch(members)
    ng, mg = b.groups()
    _say("groups", ng, mg, "graph_steps", graph_steps)
    assert ng == 3 and mg[0] == mg[1] and mg[2] == mg[3] and len({mg[0], mg[2], mg[4]}) == 3
    b.graph(graph_steps)
    b.rk4_steps(dt, E)
    tr, herm = b.check()
    _say("monitors", tr, herm)
    assert tr < 1e-10
    for i, m in enumerate(members):
        b.store(i)
        got = m.get_state()
        diff = np.abs(got - alone[i]).max()
        _say("member", i, MEMBERS[i], "max|batch - alone|", diff)
        assert np.array_equal(got, alone[i]), i
        pb, pw, v = b.observables(i)
        pb2, pw2, v2 = m.observables()
        assert np.allclose(pb, pb2, rtol=0, atol=1e-14) and np.allclose(v, v2, rtol=0, atol=1e-14)
    b.close()


def test_batch_repeated_calls_and_load(model_dir):
    """Two calls (graph replays plus leftover steps), a member state loaded
    in between: the batch follows the member contexts step for step."""
    setups = _setups(model_dir)[:2]
    n, dt = 2, 0.05
    E = _fields(14, n, dt, 6)
    members = [_engine(p, o) for p, o in setups]
    ref = [_engine(p, o) for p, o in setups]
    b = edus.Batch(members)
    b.graph(3)
    b.rk4_steps(dt, E[:7])
    for i in range(n):
        ref[i].rk4_steps(dt, E[:7, :, i, :])
    # restart member 1 from a perturbed state in both
    rho = ref[1].get_state()
    rho = rho + 1e-3 * np.eye(rho.shape[1])[None]
    ref[1].set_state(rho)
    members[1].set_state(rho)
    b.load(1)
    b.rk4_steps(dt, E[7:])
    for i in range(n):
        ref[i].rk4_steps(dt, E[7:, :, i, :])
        b.store(i)
        assert np.array_equal(members[i].get_state(), ref[i].get_state()), i
    b.close()


@pytest.mark.parametrize("idx", [0, 2, 4])
def test_context_graph_replay_and_clone(model_dir, idx):
    """edus_graph on a single context and edus_clone: same states as direct
    launches of the original context."""
    p, o = _setups(model_dir)[idx]
    dt = 0.05
    E = _fields(13, 1, dt, 7)[:, :, 0, :]
    direct = _engine(p, o)
    replay = direct.clone()
    replay.graph(5)  # two replays of 5 steps, 3 direct steps
    direct.rk4_steps(dt, E)
    replay.rk4_steps(dt, E)
    a, b = direct.get_state(), replay.get_state()
    _say(MEMBERS[idx], "max|graph - direct|", np.abs(a - b).max())
    assert np.array_equal(a, b)
    tr, _ = replay.check()
    assert tr < 1e-10


def test_batch_groups_split_by_gradient_space_and_engine(model_dir):
    """Members of one shape that differ in gradient space (spectral R vs
    finite-difference k), in the forced commutator engine or in the storage
    the commutator reads H0 and xi from (upper triangles or every element)
    form separate groups, and each still ends where it ends alone."""
    p, o = _setups(model_dir)[0]
    pk = dict(p, gradient_space="k")
    ok = oracle_py.Oracle(pk)
    m_r = _engine(p, o)
    m_k = edus.Engine(ok.nb, ok.N0, ok.N1, H0k=ok.get("H0k"), xik=(ok.get("rk0"), ok.get("rk1"), ok.get("rk2")),
                      A=ok.lattice(), filledbands=pk["filledbands"], U=ok.get("U"), gradient_space="k")
    m_k.set_state(ok.rho)
    m_e = _engine(p, o)
    t = m_e.get_tuning()
    t.comm_kind = 1 - t.comm_kind  # the other FP64 engine (FMA <-> MFMA)
    t.comm_tile = t.comm_tile_ab = 0
    m_e.set_tuning(t)
    m_t = _engine(p, o)
    t = m_t.get_tuning()
    assert t.tri_fields == 1 and m_r.get_tuning().tri_fields == 1
    t.tri_fields = 0
    m_t.set_tuning(t)
    members = [m_r, m_k, m_e, m_t]
    alone = [m.clone() for m in members]
    assert alone[3].get_tuning().tri_fields == 0  # a clone keeps the choice
    n, dt = 4, 0.05
    E = _fields(6, n, dt, 9)
    b = edus.Batch(members)
    ng, mg = b.groups()
    _say("groups", ng, mg)
    assert ng == 4 and len(set(mg)) == 4
    b.rk4_steps(dt, E)
    for i in range(n):
        alone[i].rk4_steps(dt, E[:, :, i, :])
        b.store(i)
        assert np.array_equal(members[i].get_state(), alone[i].get_state()), i
    b.close()
