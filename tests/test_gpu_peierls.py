"""GPU parity of the Peierls-gauge path (Simulation.cpp:876-899 and the
"PeierlsSubstitution" branch of the source term) against the oracle, whose
Peierls branch is pinned on the CPU by test_oracle_peierls.py (grid-shift
identity, vector potential against quadrature, two-laser state).

Bar: max|d f| <= 1e-11 relative for one right-hand side, max|d rho| <= 1e-10
and observables within 1e-10 after a propagation.
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


def _model(name, model_dir, grid, **kw):
    g = [int(grid[0]), int(grid[1]), 1]
    p = models.workload(name, os.path.join(model_dir, f"{name}_{g[0]}x{g[1]}"), grid=g, **kw)
    p["peierls"] = True
    return p


def _engine(o, p, layout):
    coul = bool(p["coulomb"])
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=p["filledbands"], U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")),
                layout=layout, peierls=True)
    if coul:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    return edus.Engine(o.nb, o.N0, o.N1, **args)


def _err(f, f_ref, layout):
    if layout == FULL:
        return np.abs(f - f_ref).max()
    iu = np.triu_indices(f.shape[1])
    return np.abs(f[:, iu[0], iu[1]] - f_ref[:, iu[0], iu[1]]).max()


RHS_CASES = [
    ("ipa2_square30", [9, 7], FULL), ("ipa2_square30", [9, 7], HERM), ("ipa2_square30", [16, 12], FULL),
    ("ipa14_hex120", [8, 10], FULL), ("ipa14_hex120", [9, 9], HERM),
    ("hsex8_hex90", [9, 11], FULL), ("hsex8_hex90", [9, 11], HERM), ("hsex8_hex90", [10, 12], FULL),
    # two-step transform lengths with the phase mode: 120 = 8*15, 20 = 5*4
    ("ipa2_square30", [120, 20], FULL),
]


@pytest.mark.parametrize("name,grid,layout", RHS_CASES)
def test_rhs_peierls_matches_oracle(model_dir, name, grid, layout):
    p = _model(name, model_dir, grid)
    _say("case", name, grid, "layout", layout)
    o = oracle_py.Oracle(p)
    eng = _engine(o, p, layout)
    rng = np.random.default_rng(3)
    for E, Av in (([3e-3, -2e-3, 0.0], [0.21, -0.13, 0.0]), ([0.0, 0.0, 0.0], [1.7, 0.4, 0.0]),
                  ([0.0, 4e-3, 0.0], [0.0, 0.0, 0.0])):
        a = rng.standard_normal((o.Nk, o.nb, o.nb)) + 1j * rng.standard_normal((o.Nk, o.nb, o.nb))
        rho = 0.5 * (a + a.conj().transpose(0, 2, 1))
        if p["coulomb"]:
            rho = o.rho + 0.05 * rho
        f_ref = o.rhs_EA(E, Av, rho)
        f = eng.rhs_peierls(E, Av, rho)
        err = _err(f, f_ref, layout)
        scale = max(1.0, np.abs(f_ref).max())
        _say("  E", E, "A", Av, "max|df|", err, "scale", scale)
        assert err <= 1e-11 * scale


def test_rhs_peierls_zero_potential_is_length_gauge_commutator(model_dir):
    """A = 0: S_A is the identity, so the Peierls right-hand side is the
    length-gauge one without its E.grad rho term (checked on the GPU alone)."""
    p = _model("ipa2_square30", model_dir, [9, 7])
    o = oracle_py.Oracle(p)
    pe = _engine(o, p, FULL)
    p2 = dict(p, peierls=False)
    o2 = oracle_py.Oracle(p2)
    le = edus.Engine(o2.nb, o2.N0, o2.N1, H0k=o2.get("H0k"), xik=(o2.get("rk0"), o2.get("rk1"), o2.get("rk2")),
                     A=o2.lattice(), filledbands=p["filledbands"], U=o2.get("U"))
    rng = np.random.default_rng(5)
    a = rng.standard_normal((o.Nk, 2, 2)) + 1j * rng.standard_normal((o.Nk, 2, 2))
    rho = 0.5 * (a + a.conj().transpose(0, 2, 1))
    # with E = 0 the length gauge has no gradient term either
    f_p = pe.rhs_peierls([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], rho)
    f_l = le.rhs([0.0, 0.0, 0.0], rho)
    assert np.abs(f_p - f_l).max() <= 1e-12 * max(1.0, np.abs(f_l).max())


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
def test_hbn_peierls_rk4_against_oracle(layout):
    p = hbn_params(grid=[15, 15, 1], peierls=True)
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=layout)
    _compare(o, sim, 60)


def test_peierls_ab4_against_oracle(model_dir):
    p = _model("ipa2_square30", model_dir, [11, 13])
    p["solver"], p["order"] = "AB", 4
    p["lasers"][0]["t0"] = -2.0
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    _compare(o, sim, 50)


def test_hsex_peierls_rk4_against_oracle(model_dir):
    p = _model("hsex8_hex90", model_dir, [9, 11])
    p["solver"] = "RK"
    p["lasers"][0]["t0"] = p["lasers"][1]["t0"] = -2.0
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    _compare(o, sim, 15)


@pytest.mark.parametrize("name,grid", [("ipa14_hex120", [8, 10]), ("hsex8_hex90", [9, 11])])
def test_rhs_peierls_mixed_precision(model_dir, name, grid):
    """Mixed precision in the Peierls gauge: the phased H(k) stays FP64 in HBM
    (it is rebuilt every stage) and the commutator runs on the FP32 matrix
    cores; relative RHS error against the FP64 oracle <= 2e-6."""
    p = _model(name, model_dir, grid)
    o = oracle_py.Oracle(p)
    coul = bool(p["coulomb"])
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=p["filledbands"], U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")),
                layout=FULL, peierls=True, precision=edus.PREC_MIXED)
    if coul:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    eng = edus.Engine(o.nb, o.N0, o.N1, **args)
    assert eng.get_tuning().comm_kind == 2
    rng = np.random.default_rng(5)
    a = rng.standard_normal((o.Nk, o.nb, o.nb)) + 1j * rng.standard_normal((o.Nk, o.nb, o.nb))
    rho = o.rho + 0.1 * 0.5 * (a + a.conj().transpose(0, 2, 1))
    E, Av = [3e-3, -2e-3, 0.0], [0.21, -0.13, 0.0]
    f_ref = o.rhs_EA(E, Av, rho)
    f = eng.rhs_peierls(E, Av, rho)
    rel = np.abs(f - f_ref).max() / np.abs(f_ref).max()
    _say(name, grid, "Peierls mixed precision: max|df| / max|f| =", rel)
    assert 1e-12 < rel <= 2e-6


def test_phased_h_reuse_is_bit_identical(model_dir):
    """S_A[H0 + E.xi] is formed once per distinct (E, A): RK4 stages 2 and 3
    share it, and a step's stage 1 reuses the previous stage 4's when the
    host times agree to the bit.  A run that forms it again at every step
    (a tuning write drops the kept one) ends bit for bit where the other does."""
    p = _model("ipa14_hex120", model_dir, [8, 10])
    p["peierls"] = True
    p["lasers"][0]["t0"] = -2.0
    a = edus.Simulation(p)
    b = edus.Simulation(p)
    a.steps(12)
    for _ in range(12):
        b.engine.set_tuning(b.engine.get_tuning())
        b.steps(1)
    ra, rb = a.engine.get_state(), b.engine.get_state()
    _say("max|rho_reuse - rho_fresh| =", np.abs(ra - rb).max(), "steps", 12)
    assert np.array_equal(ra, rb)
