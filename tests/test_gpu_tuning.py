"""GPU: kernel and tile choices, and the FP32-MFMA (mixed precision) mode.

- FFT tiles of one kernel family give bit-identical right-hand sides;
- the two FP64 commutator engines (FMA, MFMA) agree to rounding and both
  match the oracle at the parity bar;
- mixed precision (FP32 MFMA commutator, FP64 state) stays within its stated
  error of the FP64 oracle; the error is printed per case.
"""
import itertools
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


def _engine(o, layout, filled, precision=edus.PREC_FP64, coulomb=False):
    args = dict(H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")), A=o.lattice(),
                filledbands=filled, U=o.get("U"), vel=(o.get("Vel0"), o.get("Vel1"), o.get("Vel2")),
                layout=layout, precision=precision)
    if coulomb:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    return edus.Engine(o.nb, o.N0, o.N1, **args)


def _rho(nk, nb, seed, hermitian):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))
    return 0.5 * (a + a.conj().transpose(0, 2, 1)) if hermitian else a


def _upper(f):
    iu = np.triu_indices(f.shape[1])
    return f[:, iu[0], iu[1]]


def _model(model_dir, name, grid):
    return models.workload(name, os.path.join(model_dir, f"{name}_{grid[0]}x{grid[1]}"), grid=[grid[0], grid[1], 1])


@pytest.mark.parametrize("name,grid,layout", [("ipa2_square30", [40, 48], FULL), ("hsex8_hex90", [45, 27], FULL),
                                              ("hsex8_hex90", [15, 21], HERM)])
def test_fft_tiles_give_identical_results(model_dir, name, grid, layout):
    p = _model(model_dir, name, grid)
    o = oracle_py.Oracle(p)
    eng = _engine(o, layout, p["filledbands"], coulomb=bool(p["coulomb"]))
    base = eng.get_tuning()
    _say("case", name, grid, layout, "autotuned", base.as_dict())
    rho = _rho(o.Nk, o.nb, 3, hermitian=(layout == HERM))
    E = [2e-3, -1e-3, 0.0]
    ref = eng.rhs(E, rho)
    tried = 0
    for cfg0, cfg1, l0, l1 in itertools.product(range(5), range(5), (1, 3, 8, 16, 24), (1, 5, 16, 32)):
        t = edus.Tuning()
        t.fft_family[:] = list(base.fft_family)
        t.fft_cfg[:] = [cfg0, cfg1]
        t.fft_lines[:] = [l0, l1]
        t.comm_kind = base.comm_kind
        t.comm_tile = base.comm_tile
        t.comm_tile_ab = base.comm_tile_ab
        t.tri_fields = base.tri_fields
        t.hf_fused = base.hf_fused
        try:
            eng.set_tuning(t)
        except edus.EdusError:
            continue  # this tile does not fit the axis
        tried += 1
        got = eng.rhs(E, rho)
        assert np.array_equal(got, ref), t.as_dict()
    _say("  tiles tried", tried)
    assert tried >= 4


@pytest.mark.parametrize("name,grid", [("ipa2_square30", [12, 15]), ("hsex8_hex90", [9, 11]),
                                       ("ipa14_hex120", [10, 9]), ("ipa14_hex120", [8, 8])])
@pytest.mark.parametrize("layout", [FULL, HERM])
def test_fp64_commutator_engines(model_dir, name, grid, layout):
    p = _model(model_dir, name, grid)
    coul = bool(p["coulomb"])
    o = oracle_py.Oracle(p)
    if layout == HERM and coul and (grid[0] % 2 == 0 or grid[1] % 2 == 0):
        # the exchange term is not Hermitian at the even-grid Nyquist R
        with pytest.raises(edus.EdusError, match="HERMITIAN layout with exchange"):
            _engine(o, layout, p["filledbands"], coulomb=coul)
        return
    eng = _engine(o, layout, p["filledbands"], coulomb=coul)
    rho = _rho(o.Nk, o.nb, 5, hermitian=(layout == HERM))
    E = [1e-3, 3e-3, 0.0]
    f_ref = o.rhs_E(E, rho)
    scale = max(1.0, np.abs(f_ref).max())
    outs = {}
    for kind, tile in ((0, 0), (1, 0), (1, 1), (1, 2)):
        t = eng.get_tuning()
        t.comm_kind = kind
        t.comm_tile = tile
        t.comm_tile_ab = tile
        eng.set_tuning(t)
        f = eng.rhs(E, rho)
        err = np.abs(f - f_ref).max() if layout == FULL else np.abs(_upper(f) - _upper(f_ref)).max()
        _say(name, grid, "layout", layout, "comm_kind", kind, "tile", tile, "max|df| vs oracle", err, "scale", scale)
        assert err <= 1e-11 * scale
        outs[(kind, tile)] = f
    assert np.abs(outs[(0, 0)] - outs[(1, 0)]).max() <= 1e-13 * scale
    assert np.array_equal(outs[(1, 0)], outs[(1, 1)])  # tiles: same arithmetic per k-point
    assert np.array_equal(outs[(1, 0)], outs[(1, 2)])


@pytest.mark.parametrize("name,grid", [("hsex8_hex90", [9, 11]), ("ipa14_hex120", [10, 9])])
@pytest.mark.parametrize("order", [4, 5])
def test_adams_bashforth_tiles_identical(model_dir, name, grid, order):
    """The Adams-Bashforth kernel (history prefetched into registers) gives
    bit-identical states with either tile (its parity with the oracle is
    checked in test_gpu_parity)."""
    p = _model(model_dir, name, grid)
    p["solver"], p["order"] = "AB", order
    o = oracle_py.Oracle(p)
    coul = bool(p["coulomb"])
    states = []
    for tile in (0, 1, 2):
        eng = _engine(o, FULL, p["filledbands"], coulomb=coul)
        eng.set_state(o.rho)
        t = eng.get_tuning()
        t.comm_kind = 1
        t.comm_tile_ab = tile
        eng.set_tuning(t)
        rng = np.random.default_rng(4)
        E = rng.uniform(-4e-3, 4e-3, size=(12, 3)) * np.array([1.0, 1.0, 0.0])
        eng.ab_steps(0.05, order, E)
        states.append(eng.get_state())
        tr, herm = eng.check()
        _say(name, grid, "AB", order, "tile", tile, "monitors", tr, herm)
    assert np.array_equal(states[0], states[1])
    assert np.array_equal(states[0], states[2])


@pytest.mark.parametrize("name,grid", [("ipa2_square30", [12, 15]), ("ipa14_hex120", [10, 9]),
                                       ("hsex8_hex90", [9, 11])])
def test_mixed_precision_rhs_error(model_dir, name, grid):
    """FP32-MFMA commutator: relative RHS error against the FP64 oracle."""
    p = _model(model_dir, name, grid)
    o = oracle_py.Oracle(p)
    eng = _engine(o, FULL, p["filledbands"], precision=edus.PREC_MIXED, coulomb=bool(p["coulomb"]))
    assert eng.get_tuning().comm_kind == 2
    rho = o.rho + 0.1 * _rho(o.Nk, o.nb, 9, hermitian=True)  # off the ground state: [H, rho] = O(|H|)
    E = [5e-3, -2e-3, 0.0]
    f_ref = o.rhs_E(E, rho)
    f = eng.rhs(E, rho)
    rel = np.abs(f - f_ref).max() / np.abs(f_ref).max()
    _say(name, grid, "mixed precision: max|df| / max|f| =", rel)
    assert rel <= 2e-6
    for tile in (0, 1, 2):  # the FP32 engine's tiles: same arithmetic per k-point
        t = eng.get_tuning()
        t.comm_tile = tile
        eng.set_tuning(t)
        assert np.array_equal(eng.rhs(E, rho), f)


def test_mixed_precision_propagation_error():
    """hBN, 400 RK4 steps in mixed precision against the FP64 oracle:
    populations and current, and the density matrix itself."""
    p = hbn_params()
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, precision=edus.PREC_MIXED)
    ref = o.propagate(400)
    got = sim.propagate(400)
    drho = np.abs(sim.engine.get_state() - o.rho).max()
    dpop = np.abs(got["pop_bloch"] - ref["pop_bloch"]).max()
    vref = np.abs(ref["velocity"]).max()
    dvel = np.abs(got["velocity"] - ref["velocity"]).max() / vref
    _say("mixed precision hBN 400 steps: max|d rho|", drho, "max|d pop|", dpop, "rel d current", dvel)
    assert drho <= 1e-5
    assert dpop <= 1e-5
    assert dvel <= 1e-3


@pytest.mark.parametrize("grid", [(120, 120), (96, 96), (90, 90), (48, 48)])
def test_fused_gradient_identical_to_two_pass(model_dir, grid):
    """The one-pass spectral derivative (one work-group per plane) against the
    two axis passes and the oracle: the same DFT arithmetic in another
    kernel, so the two forms agree to rounding (the compiler contracts
    multiply-adds per kernel); the field has components along both lattice
    vectors so both terms act."""
    p = _model(model_dir, "ipa2_square30", list(grid))
    o = oracle_py.Oracle(p)
    eng = _engine(o, FULL, p["filledbands"])
    rho = _rho(o.Nk, o.nb, 11, hermitian=False)
    E = [3e-3, -2e-3, 0.0]
    f_ref = o.rhs_E(E, rho)
    outs = []
    for fused in (0, 1):
        t = eng.get_tuning()
        t.grad_fused = fused
        eng.set_tuning(t)
        outs.append(eng.rhs(E, rho))
    err = np.abs(outs[1] - f_ref).max()
    d = np.abs(outs[1] - outs[0]).max()
    scale = max(1.0, np.abs(f_ref).max())
    _say(grid, "fused gradient: max|df| vs oracle", err, "vs two passes", d, "scale", scale)
    assert d <= 1e-14 * scale
    assert err <= 1e-11 * scale


def test_fused_gradient_with_field_tables(model_dir):
    """Graph replay reads the field and the gradient coefficients from device
    tables (the batch path): the one-pass gradient with those tables ends
    where its direct launches end, bit for bit, and within rounding of the
    two-pass gradient."""
    p = _model(model_dir, "ipa2_square30", [96, 96])
    o = oracle_py.Oracle(p)
    runs = []
    for fused, graph in ((1, 0), (1, 5), (0, 0)):
        eng = _engine(o, FULL, p["filledbands"])
        t = eng.get_tuning()
        t.grad_fused = fused
        eng.set_tuning(t)
        eng.graph(graph)
        rng = np.random.default_rng(8)
        E = rng.uniform(-4e-3, 4e-3, size=(10, 3, 3)) * np.array([1.0, 1.0, 0.0])
        eng.rk4_steps(0.05, E)
        runs.append(eng.get_state())
    assert np.abs(runs[0] - o.rho).max() > 0  # the field moved the state
    assert np.array_equal(runs[0], runs[1])
    _say("fused vs two-pass after 10 RK4 steps:", np.abs(runs[0] - runs[2]).max())
    assert np.abs(runs[0] - runs[2]).max() <= 1e-14


def test_set_tuning_after_graph_replay(model_dir):
    """edus_set_tuning drops a captured graph: after graph replays, forcing
    the other FP64 commutator engine takes effect in the next replays, which
    then end where direct launches with that engine end."""
    p = _model(model_dir, "ipa14_hex120", [10, 9])
    o = oracle_py.Oracle(p)
    rng = np.random.default_rng(12)
    E = rng.uniform(-4e-3, 4e-3, size=(10, 3, 3)) * np.array([1.0, 1.0, 0.0])
    states = []
    for replay_after in (True, False):
        eng = _engine(o, FULL, p["filledbands"])
        eng.set_state(o.rho)
        eng.graph(5)
        eng.rk4_steps(0.05, E[:5])
        t = eng.get_tuning()
        t.comm_kind = 1 - t.comm_kind
        t.comm_tile = t.comm_tile_ab = 0
        eng.set_tuning(t)
        eng.graph(5 if replay_after else 0)
        eng.rk4_steps(0.05, E[5:])
        states.append(eng.get_state())
    assert np.array_equal(states[0], states[1])


@pytest.mark.parametrize("grid,layout", [([9, 11], FULL), ([9, 30], FULL), ([9, 15], HERM), ([10, 24], FULL)])
def test_exchange_row_pass_fused_identical(model_dir, grid, layout):
    """With exchange, the gradient's row pass can also store the exchange
    term's forward row transform (hf_fused 1, the default): right-hand sides
    and RK4 / AB4 states equal to the separate pass's to rounding (the same
    DFT arithmetic in another kernel; the compiler contracts multiply-adds per
    kernel, and most shapes come out bit-identical), with the field along
    both grid axes, along axis 1 only and along axis 0 only (no row pass:
    nothing to fuse); row lengths with one and with two register steps."""
    p = _model(model_dir, "hsex8_hex90", grid)
    o = oracle_py.Oracle(p)
    eng = _engine(o, layout, p["filledbands"], coulomb=True)
    t = eng.get_tuning()
    assert t.hf_fused == 1 and t.fft_family[1] == 1
    rho = o.rho + 0.05 * _rho(o.Nk, o.nb, 4, hermitian=True)
    A = o.lattice()
    # c_j = a_j . E: a field perpendicular to a_1 (a_0) has no axis-1 (axis-0) term
    z = np.array([0.0, 0.0, 1.0])
    e0, e1 = np.cross(A[:, 1], z), np.cross(A[:, 0], z)
    fields = {"both": [3e-3, -2e-3, 0.0], "axis 0 only": list(3e-3 * e0 / np.linalg.norm(e0)),
              "axis 1 only": list(3e-3 * e1 / np.linalg.norm(e1))}
    for what, E in fields.items():
        f = {}
        for on in (1, 0):
            t = eng.get_tuning()
            t.hf_fused = on
            eng.set_tuning(t)
            f[on] = eng.rhs(E, rho)
        _say(grid, layout, what, "fused vs separate: max|df|", np.abs(f[1] - f[0]).max())
        assert np.abs(f[1] - f[0]).max() <= 1e-15 * np.abs(f[0]).max()
        if what == "both":
            ref = o.rhs_E(E, rho)
            err = np.abs(f[1] - ref).max() if layout == FULL else np.abs(_upper(f[1]) - _upper(ref)).max()
            assert err <= 1e-11 * max(1.0, np.abs(ref).max())
    rng = np.random.default_rng(2)
    Es = rng.uniform(-4e-3, 4e-3, size=(6, 3, 3)) * np.array([1.0, 1.0, 0.0])
    Eab = rng.uniform(-4e-3, 4e-3, size=(6, 3)) * np.array([1.0, 1.0, 0.0])
    states = {}
    for on in (1, 0):
        t = eng.get_tuning()
        t.hf_fused = on
        eng.set_tuning(t)
        eng.set_state(o.rho)
        eng.ab_reset()
        eng.rk4_steps(0.05, Es)
        eng.ab_steps(0.05, 4, Eab)
        states[on] = eng.get_state()
    _say(grid, layout, "6 RK4 + 6 AB4 steps, fused vs separate: max|d rho|", np.abs(states[1] - states[0]).max())
    assert np.abs(states[1] - states[0]).max() <= 1e-14
