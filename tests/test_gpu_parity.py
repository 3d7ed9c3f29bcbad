"""GPU parity: the HIP engine (through the C ABI) against the FP64 CPU oracle
and the reference's committed golden outputs.

Bar: max|d rho| <= 1e-10 and observables within 1e-10 absolute
(ci-test/compare.py), relative current error <= 1e-9.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, hbn_params

import edus
from edus import models

oracle_py = pytest.importorskip("oracle_py")

pytestmark = pytest.mark.gpu

FULL, HERM = edus.LAYOUT_FULL, edus.LAYOUT_HERMITIAN


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine_from_oracle(o, layout, filled, coulomb=False, rho0=None):
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=filled, U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")), layout=layout)
    if coulomb:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=rho0)
    return edus.Engine(o.nb, o.N0, o.N1, **args)


def _rhs_err(f, f_ref, layout):
    """FULL: every element.  HERMITIAN: the stored upper triangle (incl. the
    diagonal) -- the lower one is its conjugate by construction, while the
    reference's lower triangle carries the even-grid Nyquist term."""
    if layout == FULL:
        return np.abs(f - f_ref).max()
    iu = np.triu_indices(f.shape[1])
    return np.abs(f[:, iu[0], iu[1]] - f_ref[:, iu[0], iu[1]]).max()


def random_hermitian(nk, nb, seed, hermitian=True):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))
    if hermitian:
        a = 0.5 * (a + a.conj().transpose(0, 2, 1))
    return a


def _model(name, model_dir, grid, **kw):
    g = [int(grid[0]), int(grid[1]), 1]
    return models.workload(name, os.path.join(model_dir, f"{name}_{g[0]}x{g[1]}"), grid=g, **kw)


RHS_CASES = [
    ("ipa2_square30", [6, 6]), ("ipa2_square30", [9, 7]), ("ipa2_square30", [30, 30]),
    ("ipa14_hex120", [6, 8]), ("ipa14_hex120", [10, 9]), ("hsex8_hex90", [12, 12]), ("hsex8_hex90", [9, 11]),
    # the benchmark grids' axis lengths: two-step kernels 120 = 8*15, 96 = 8*12,
    # 90 = 10*9, 48 = 8*6, 20 = 5*4 (every transform mode with exchange), and
    # 45, 27 on the Stockham kernel
    ("ipa2_square30", [120, 120]), ("ipa2_square30", [96, 90]), ("ipa2_square30", [48, 20]),
    ("hsex8_hex90", [90, 96]), ("hsex8_hex90", [45, 27]),
]


@pytest.mark.parametrize("layout", [FULL, HERM])
@pytest.mark.parametrize("name,grid", RHS_CASES)
def test_rhs_matches_oracle(model_dir, name, grid, layout):
    p = _model(name, model_dir, grid)
    _say("case", name, grid, "layout", layout)
    coul = bool(p["coulomb"])
    o = oracle_py.Oracle(p)
    if layout == HERM and coul and (grid[0] % 2 == 0 or grid[1] % 2 == 0):
        # on even grids the reference evaluates W at the unpaired Nyquist R, so
        # its exchange term is not Hermitian there: the engine refuses the layout
        with pytest.raises(edus.EdusError, match="HERMITIAN layout with exchange"):
            engine_from_oracle(o, layout, p["filledbands"], coulomb=coul, rho0=o.rho)
        return
    eng = engine_from_oracle(o, layout, p["filledbands"], coulomb=coul, rho0=o.rho)
    for seed, E in ((1, [3e-3, -2e-3, 0.0]), (2, [0.0, 4e-3, 0.0]), (3, [0.0, 0.0, 0.0])):
        rho = random_hermitian(o.Nk, o.nb, seed, hermitian=(layout == HERM))
        f_ref = o.rhs_E(E, rho)
        f = eng.rhs(E, rho)
        err = _rhs_err(f, f_ref, layout)
        scale = max(1.0, np.abs(f_ref).max())
        _say("  E", E, "max|df|", err, "scale", scale)
        assert err <= 1e-11 * scale


@pytest.mark.parametrize("layout,n", [(FULL, 20), (HERM, 21)])
def test_rhs_hbn_coulomb_from_oracle_arrays(layout, n):
    """hBN with Hartree + screened exchange: engine kernels fed with the
    oracle's own set-up arrays (isolates the kernels from the host set-up)."""
    p = hbn_params(coulomb=True, grid=[n, n, 1])
    o = oracle_py.Oracle(p)
    eng = engine_from_oracle(o, layout, 1, coulomb=True, rho0=o.rho)
    rng = np.random.default_rng(7)
    for E in ([0.0, 0.0, 0.0], [1e-3, 2e-3, 0.0]):
        d = 0.01 * (rng.standard_normal((o.Nk, 2, 2)) + 1j * rng.standard_normal((o.Nk, 2, 2)))
        rho = o.rho + 0.5 * (d + d.conj().transpose(0, 2, 1))
        f_ref = o.rhs_E(E, rho)
        f = eng.rhs(E, rho)
        err = _rhs_err(f, f_ref, layout)
        _say("E", E, "max|df|", err, "scale", np.abs(f_ref).max())
        assert err <= 1e-12


def _compare_runs(o, sim, nsteps):
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
    return got


def test_hbn_golden_on_gpu():
    """The reference's own hBN integration test, end to end through the engine."""
    p = hbn_params()
    sim = edus.Simulation(p)
    got = sim.propagate()
    assert got["pop_bloch"].shape == (102, 2)
    for name, key in (("Population.txt", "pop_bloch"), ("Velocity.txt", "velocity"), ("Laser.txt", "laser")):
        gold = np.loadtxt(os.path.join(GOLDEN, "hBN", name))
        diff = np.abs(got[key] - gold)
        bad = (diff > 1e-10) & ((np.abs(got[key]) > 1e-10) | (np.abs(gold) > 1e-10))
        _say(name, "max diff", diff.max())
        assert not bad.any(), name
    tr, herm = sim.engine.check()
    _say("monitors", tr, herm)
    assert tr < 1e-12


@pytest.mark.parametrize("layout", [FULL, HERM])
def test_hbn_against_oracle(layout):
    p = hbn_params()
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=layout)
    _compare_runs(o, sim, 60)


PROP_CASES = [
    ("ipa2_square30", [15, 15], FULL, 80), ("ipa2_square30", [15, 15], HERM, 80),
    ("ipa2_square30", [16, 12], FULL, 80),
    ("ipa14_hex120", [9, 9], FULL, 20), ("ipa14_hex120", [9, 9], HERM, 20),
    ("ipa14_hex120", [8, 10], FULL, 20),
]


@pytest.mark.parametrize("name,grid,layout,nsteps", PROP_CASES)
def test_rk4_propagation_matches_oracle(model_dir, name, grid, layout, nsteps):
    p = _model(name, model_dir, grid)
    # start inside the pulse so the field is large from the first step
    p["lasers"][0]["t0"] = -2.0
    _say("case", name, grid, layout, nsteps)
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=layout)
    _compare_runs(o, sim, nsteps)
    tr, herm = sim.engine.check()
    _say("  monitors", tr, herm)
    assert tr < 1e-11
    if layout == HERM or grid[0] % 2 == 1:
        assert herm < 1e-12


@pytest.mark.parametrize("order", [4, 5])
@pytest.mark.parametrize("layout", [FULL, HERM])
def test_adams_bashforth_matches_oracle(model_dir, order, layout):
    p = _model("ipa2_square30", model_dir, [11, 13])
    p["solver"], p["order"] = "AB", order
    p["lasers"][0]["t0"] = -2.0
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=layout)
    _compare_runs(o, sim, 60)


@pytest.mark.parametrize("layout", [FULL, HERM])
def test_hsex_rk4_matches_oracle(model_dir, layout):
    p = _model("hsex8_hex90", model_dir, [9, 11])
    p["solver"] = "RK"
    p["lasers"][0]["t0"] = p["lasers"][1]["t0"] = -2.0
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=layout)
    _compare_runs(o, sim, 15)


def test_hsex_ab4_even_grid_full(model_dir):
    p = _model("hsex8_hex90", model_dir, [10, 10])
    p["lasers"][0]["t0"] = p["lasers"][1]["t0"] = -2.0
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=FULL)
    _compare_runs(o, sim, 25)


def test_hbn_coulomb_with_reference_quirk():
    """hBN with HSEX and the reference's own centre accumulation
    (ModelCoulomb.cpp:76-79): harmless for hBN because centre 0 is at the origin."""
    p = hbn_params(coulomb=True, grid=[20, 20, 1])
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=FULL)
    _compare_runs(o, sim, 30)


def test_hbn_coulomb_reference_fixture_low_level():
    """The reference's own HSEX run (ci-test/outputs/hBN_coulomb): the fixture
    was written by a code version that read W(R) from RytovaKeldysh.py's file
    (committed beside it) and had no Hartree term.  Driving the engine through
    the low-level ABI with that W reproduces the reference's Bloch populations
    to 1e-10 (compare.py bar) and the current to 1e-9."""
    p = hbn_params(coulomb=True)
    o = oracle_py.Oracle(p)
    W = np.loadtxt(os.path.join(GOLDEN, "hBN_coulomb", "RytovaKeldysh.txt"), skiprows=1).reshape(-1, 2, 2)
    eng = edus.Engine(o.nb, o.N0, o.N1, H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")),
                      A=o.lattice(), filledbands=1, U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")),
                      coulomb=True, W=W.astype(complex), hartree=np.zeros((2, 2), complex), rho0k=o.rho)
    eng.set_state(o.rho)
    gold_p = np.loadtxt(os.path.join(GOLDEN, "hBN_coulomb", "Population.txt"))
    gold_v = np.loadtxt(os.path.join(GOLDEN, "hBN_coulomb", "Velocity.txt"))
    t, dt = 0.0, o.dt
    pops, vels = [], []
    for _ in range(o.nsteps):
        pb, _pw, v = eng.observables()
        pops.append(pb)
        vels.append(v)
        eng.rk4_steps(dt, np.array([o.E(t), o.E(t + dt / 2.0), o.E(t + dt)]))
        t += dt
    dp = np.abs(np.array(pops) - gold_p).max()
    dv = np.abs(np.array(vels) - gold_v).max()
    _say("hBN_coulomb fixture: max|dP|", dp, "max|dJ|", dv)
    assert dp <= 1e-10
    assert dv <= 1e-9


def test_setup_arrays_match_oracle(model_dir):
    """Host set-up (DFT, make_hermitian, eigenvalues, velocity operator, W, Hartree)."""
    p = _model("hsex8_hex90", model_dir, [8, 9])
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, layout=FULL)
    for name, oname in (("H0k", "H0k"), ("rk0", "rk0"), ("rk1", "rk1"), ("W", "Wscr"), ("hartree", "hartree"),
                        ("Vel0", "Vel0"), ("Vel1", "Vel1"), ("Vel2", "Vel2")):
        a, b = sim.get(name), o.get(oname)
        err = float(np.abs(a - b).max())
        _say(name, err)
        assert err <= 1e-11 * max(1.0, float(np.abs(b).max())), name
    assert np.abs(sim.get("eps") - o.eigenvalues()).max() < 1e-12


def test_error_paths():
    p = hbn_params()
    sim = edus.Simulation(p)
    with pytest.raises(edus.EdusError) as e:
        sim.engine.ab_steps(0.1, 3, np.zeros((1, 3)))
    assert e.value.code == 4
    bad = dict(p)
    bad["tb_file"] = "/nonexistent/model"
    with pytest.raises(edus.EdusError):
        edus.Simulation(bad)


def test_hbn_reference_output_files(tmp_path):
    """The reference's output files written by the library, compared with its
    own hBN fixtures column by column at compare.py's 1e-10."""
    sim = edus.Simulation(hbn_params())
    k = sim.write_outputs(str(tmp_path))
    assert k == 102
    for name in ("Time.txt", "Laser.txt", "Population.txt", "Velocity.txt"):
        got = np.loadtxt(os.path.join(str(tmp_path), name))
        gold = np.loadtxt(os.path.join(GOLDEN, "hBN", name))
        assert got.shape == gold.shape, name
        tol = 1e-10 if name != "Time.txt" else 5e-6 * np.abs(gold).max()  # Time.txt: 6 printed digits
        _say(name, "max diff", np.abs(got - gold).max())
        assert np.abs(got - gold).max() <= tol, name
    assert np.loadtxt(os.path.join(str(tmp_path), "Population_wannier.txt")).shape == (102, 2)
    # DM0.txt: distance |r_a + R - r_b| and rho(R) at t = 0, against the reference's file
    got = np.loadtxt(os.path.join(str(tmp_path), "DM0.txt"))
    gold = np.loadtxt(os.path.join(GOLDEN, "hBN", "DM0.txt"))
    assert got.shape == gold.shape
    _say("DM0.txt max diff", np.abs(got - gold).max(axis=0))
    assert np.abs(got[:, 0] - gold[:, 0]).max() <= 1e-8 * np.abs(gold[:, 0]).max()  # 10 printed digits
    assert np.abs(got[:, 1:] - gold[:, 1:]).max() <= 1e-10
    # Laser_A.txt: the vector potential evaluated at every print (no reference
    # fixture; against the oracle, whose integration is pinned by quadrature)
    o = oracle_py.Oracle(hbn_params())
    ref = o.propagate()
    la = np.loadtxt(os.path.join(str(tmp_path), "Laser_A.txt"))
    _say("Laser_A.txt max diff", np.abs(la - ref["laser_A"]).max(), "max |A|", np.abs(ref["laser_A"]).max())
    assert la.shape == ref["laser_A"].shape
    assert np.abs(la - ref["laser_A"]).max() <= 1e-12 * max(1.0, np.abs(ref["laser_A"]).max())


def test_ab4_seeded_by_rk4(model_dir):
    """edus_ab_seed: the start-up steps are RK4 steps (bit for bit), and the
    seeded AB4 run stays closer to fine RK4 than the reference's zero-history
    start."""
    p = _model("ipa2_square30", model_dir, [11, 13])
    p["lasers"][0]["t0"] = -2.0
    dt = 0.05
    sim = edus.Simulation(p)
    e0 = sim.engine
    E9 = np.array([[sim.field(dt * (s + f)) for f in (0.0, 0.5, 1.0)] for s in range(400)])  # [400][3][3]
    ref, seeded, zero = e0.clone(), e0.clone(), e0.clone()
    rk3 = e0.clone()
    rk3.rk4_steps(dt, E9[:3])
    seeded.ab_seed(dt, 4, E9[:3])
    assert np.array_equal(seeded.get_state(), rk3.get_state())
    n = 200
    seeded.ab_steps(dt, 4, E9[3:3 + n, 0, :])
    zero.ab_steps(dt, 4, E9[:3 + n, 0, :])
    ref.rk4_steps(dt, E9[:3 + n])
    es = np.abs(seeded.get_state() - ref.get_state()).max()
    ez = np.abs(zero.get_state() - ref.get_state()).max()
    _say("AB4 vs RK4 after", 3 + n, "steps: seeded", es, "zero-history", ez)
    assert es < ez
