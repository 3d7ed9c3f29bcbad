"""GPU parity of gradient_space "k" (the finite-difference k gradient,
kGradient.cpp:25-210, kGradient_definitions.hpp:19-31) against the oracle,
whose stencil is pinned on the CPU by test_oracle_kgradient.py (B1 condition,
central differences, second-order convergence on a plane wave) and matches
the library's host-side builder there.

Bar: one RHS within 1e-11 relative; propagations max|d rho| <= 1e-10 and
observables within 1e-10; batches equal single contexts bit for bit.
"""
import os
import sys

import numpy as np
import pytest

from conftest import hbn_params

import edus
from edus import models

oracle_py = pytest.importorskip("oracle_py")

pytestmark = pytest.mark.gpu

FULL, HERM = edus.LAYOUT_FULL, edus.LAYOUT_HERMITIAN


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _model(name, model_dir, grid):
    g = [int(grid[0]), int(grid[1]), 1]
    p = models.workload(name, os.path.join(model_dir, f"{name}_{g[0]}x{g[1]}"), grid=g)
    p["gradient_space"] = "k"
    return p


def _engine(o, p, layout):
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=p["filledbands"], U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")),
                layout=layout, gradient_space="k")
    if p["coulomb"]:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    return edus.Engine(o.nb, o.N0, o.N1, **args)


def _err(f, f_ref, layout):
    if layout == FULL:
        return np.abs(f - f_ref).max()
    iu = np.triu_indices(f.shape[1])
    return np.abs(f[:, iu[0], iu[1]] - f_ref[:, iu[0], iu[1]]).max()


RHS_CASES = [
    ("ipa2_square30", [8, 12], FULL), ("ipa2_square30", [8, 12], HERM), ("ipa2_square30", [9, 9], FULL),
    ("ipa14_hex120", [10, 9], FULL), ("ipa14_hex120", [10, 9], HERM),
    ("hsex8_hex90", [9, 11], FULL), ("hsex8_hex90", [9, 11], HERM), ("hsex8_hex90", [10, 12], FULL),
    ("ipa2_square30", [120, 96], FULL),
]


@pytest.mark.parametrize("name,grid,layout", RHS_CASES)
def test_rhs_kgradient_matches_oracle(model_dir, name, grid, layout):
    p = _model(name, model_dir, grid)
    _say("case", name, grid, "layout", layout)
    o = oracle_py.Oracle(p)
    eng = _engine(o, p, layout)
    rng = np.random.default_rng(11)
    for E in ([3e-3, -2e-3, 0.0], [0.0, 4e-3, 0.0], [0.0, 0.0, 0.0]):
        a = rng.standard_normal((o.Nk, o.nb, o.nb)) + 1j * rng.standard_normal((o.Nk, o.nb, o.nb))
        rho = 0.5 * (a + a.conj().transpose(0, 2, 1))
        if p["coulomb"]:
            rho = o.rho + 0.05 * rho
        f_ref = o.rhs_E(E, rho)
        f = eng.rhs(E, rho)
        err = _err(f, f_ref, layout)
        scale = max(1.0, np.abs(f_ref).max())
        _say("  E", E, "max|df|", err, "scale", scale)
        assert err <= 1e-11 * scale


def test_velocity_operator_uses_the_stencil(model_dir):
    """V_d = (-i[r_d, H0] + grad_k,d H0)/det with the finite-difference gradient
    (Simulation.cpp:538-575), built by the library's simulation layer: the
    current of a perturbed state against the oracle's."""
    p = _model("ipa2_square30", model_dir, [8, 12])
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    rng = np.random.default_rng(8)
    a = rng.standard_normal((o.Nk, 2, 2)) + 1j * rng.standard_normal((o.Nk, 2, 2))
    rho = o.rho + 0.05 * 0.5 * (a + a.conj().transpose(0, 2, 1))
    o.set("rho", rho)
    sim.engine.set_state(rho)
    _, _, v = sim.engine.observables()
    _, _, v_ref = o.observables()
    _say("velocity", v, "oracle", v_ref)
    assert np.abs(v_ref).max() > 1e-6
    assert np.abs(v - v_ref).max() <= 1e-9 * np.abs(v_ref).max()


def _compare(o, sim, nsteps):
    ref = o.propagate(nsteps)
    got = sim.propagate(nsteps)
    for key in ("time", "laser", "pop_bloch", "pop_wannier", "velocity"):
        err = np.abs(got[key] - ref[key]).max()
        _say("  ", key, "max abs err", err)
        assert err <= 1e-10, key
    vref = np.abs(ref["velocity"]).max()
    if vref > 0:
        assert np.abs(got["velocity"] - ref["velocity"]).max() <= 1e-9 * vref + 1e-14
    drho = np.abs(sim.engine.get_state() - o.rho).max()
    _say("   max|d rho|", drho)
    assert drho <= 1e-10


@pytest.mark.parametrize("layout", [FULL, HERM])
def test_hbn_kgradient_rk4_against_oracle(layout):
    p = hbn_params(grid=[15, 15, 1], gradient_space="k")
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=layout)
    _compare(o, sim, 60)
    tr, herm = sim.engine.check()
    _say("  monitors", tr, herm)
    assert tr < 1e-11 and herm < 1e-12


def test_kgradient_ab4_even_grid(model_dir):
    p = _model("ipa2_square30", model_dir, [12, 10])
    p["solver"], p["order"] = "AB", 4
    p["lasers"][0]["t0"] = -2.0
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    _compare(o, sim, 50)


def test_hsex_kgradient_rk4(model_dir):
    p = _model("hsex8_hex90", model_dir, [9, 11])
    p["solver"] = "RK"
    p["lasers"][0]["t0"] = p["lasers"][1]["t0"] = -2.0
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    _compare(o, sim, 15)


@pytest.mark.parametrize("graph_steps", [0, 3])
def test_kgradient_batch_equals_single(model_dir, graph_steps):
    """Two k-gradient contexts of one lattice as a batch group, with direct
    launches and graph replay: bit-identical to each alone."""
    p = _model("ipa2_square30", model_dir, [12, 10])
    o = oracle_py.Oracle(p)
    rng = np.random.default_rng(2)
    E = np.zeros((6, 3, 2, 3))
    E[..., :2] = rng.uniform(-4e-3, 4e-3, size=(6, 3, 2, 2))
    alone = []
    for i in range(2):
        e = _engine(o, p, FULL)
        e.set_state(o.rho)
        e.rk4_steps(0.05, E[:, :, i, :])
        alone.append(e.get_state())
    members = []
    for i in range(2):
        e = _engine(o, p, FULL)
        e.set_state(o.rho)
        members.append(e)
    b = edus.Batch(members)
    b.graph(graph_steps)
    b.rk4_steps(0.05, E)
    for i, m in enumerate(members):
        b.store(i)
        assert np.array_equal(m.get_state(), alone[i]), i


@pytest.mark.parametrize("name,grid", [("ipa14_hex120", [10, 9]), ("hsex8_hex90", [9, 11])])
def test_rhs_kgradient_mixed_precision(model_dir, name, grid):
    """Mixed precision with the finite-difference k gradient: H0 and xi from
    the FP32 copies (IPA) or FP64 (exchange), the commutator on the FP32
    matrix cores, the stencil FP64; relative RHS error against the FP64
    oracle <= 2e-6."""
    p = _model(name, model_dir, grid)
    o = oracle_py.Oracle(p)
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=p["filledbands"], U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")),
                layout=FULL, gradient_space="k", precision=edus.PREC_MIXED)
    if p["coulomb"]:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    eng = edus.Engine(o.nb, o.N0, o.N1, **args)
    assert eng.get_tuning().comm_kind == 2
    rng = np.random.default_rng(6)
    a = rng.standard_normal((o.Nk, o.nb, o.nb)) + 1j * rng.standard_normal((o.Nk, o.nb, o.nb))
    rho = o.rho + 0.1 * 0.5 * (a + a.conj().transpose(0, 2, 1))
    E = [5e-3, -2e-3, 0.0]
    f_ref = o.rhs_E(E, rho)
    f = eng.rhs(E, rho)
    rel = np.abs(f - f_ref).max() / np.abs(f_ref).max()
    _say(name, grid, "k-gradient mixed precision: max|df| / max|f| =", rel)
    assert 1e-12 < rel <= 2e-6
