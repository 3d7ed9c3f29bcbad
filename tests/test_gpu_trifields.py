"""GPU: the commutator reading H0(k) and r_d(k) from their upper triangles
(tuning field tri_fields; FULL layout, Hermitian fields).

- exactly Hermitian fields: tri_fields 1 and 0 give bit-identical right-hand
  sides and RK4 / Adams-Bashforth states (both FP64 engines, and the FP32
  engine of mixed precision), since conjugating the upper element and every
  addend gives exactly the lower element's sum;
- fields from the DFT of a Hermitian H(R), Hermitian to 1-2 ulp: the two
  storages agree to a few ulp of max|f|, and both meet the oracle bar;
- where it does not apply (HERMITIAN layout, exchange, a non-Hermitian H0)
  tri_fields is 0 and asking for 1 fails.
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


def _model(model_dir, name, grid):
    return models.workload(name, os.path.join(model_dir, f"{name}_{grid[0]}x{grid[1]}"), grid=[grid[0], grid[1], 1])


def _herm(f):
    """(f + f^dagger) / 2 per k: exactly Hermitian in floating point (the two
    mirror sums add the same two numbers)."""
    return 0.5 * (f + f.conj().transpose(0, 2, 1))


def _engine(o, layout=FULL, precision=edus.PREC_FP64, exact=False, H0k=None, coulomb=False):
    H0 = o.get("H0k") if H0k is None else H0k
    xi = tuple(o.get(f"rk{d}") for d in range(3))
    if exact:
        H0 = _herm(H0)
        xi = tuple(_herm(x) for x in xi)
    args = dict(H0k=H0, xik=xi, A=o.lattice(), filledbands=o.nb // 2, U=o.get("U"), layout=layout,
                precision=precision)
    if coulomb:
        args.update(coulomb=True, W=o.get("Wscr"), hartree=o.get("hartree"), rho0k=o.rho)
    return edus.Engine(o.nb, o.N0, o.N1, **args)


def _set_tri(eng, on):
    t = eng.get_tuning()
    t.tri_fields = 1 if on else 0
    eng.set_tuning(t)


def _rho(nk, nb, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))
    return 0.5 * (a + a.conj().transpose(0, 2, 1))


@pytest.mark.parametrize("name,grid", [("ipa14_hex120", [10, 9]), ("ipa2_square30", [12, 15]),
                                       ("ipa6_hex96", [16, 16])])
@pytest.mark.parametrize("precision,kinds", [(edus.PREC_FP64, (0, 1)), (edus.PREC_MIXED, (2,))])
def test_exactly_hermitian_fields_identical(model_dir, name, grid, precision, kinds):
    p = _model(model_dir, name, grid)
    o = oracle_py.Oracle(p)
    eng = _engine(o, precision=precision, exact=True)
    assert eng.get_tuning().tri_fields == 1
    rho = o.rho + 0.1 * _rho(o.Nk, o.nb, 3)
    E = [4e-3, -3e-3, 0.0]
    for kind in kinds:
        for tile in (0, 1, 2) if kind else (0,):
            t = eng.get_tuning()
            t.comm_kind, t.comm_tile, t.comm_tile_ab = kind, tile, tile
            eng.set_tuning(t)
            f = {}
            for on in (1, 0):
                _set_tri(eng, on)
                f[on] = eng.rhs(E, rho)
            _say(name, grid, "precision", precision, "kind", kind, "tile", tile, "tri 1 vs 0: max|df|",
                 np.abs(f[1] - f[0]).max())
            assert np.array_equal(f[1], f[0])
    rng = np.random.default_rng(7)
    Es = rng.uniform(-5e-3, 5e-3, size=(8, 3, 3)) * np.array([1.0, 1.0, 0.0])
    Eab = rng.uniform(-5e-3, 5e-3, size=(8, 3)) * np.array([1.0, 1.0, 0.0])
    states = {}
    for on in (1, 0):
        _set_tri(eng, on)
        eng.set_state(o.rho)
        eng.ab_reset()
        eng.rk4_steps(0.05, Es)
        eng.ab_steps(0.05, 4, Eab)
        states[on] = eng.get_state()
    _say(name, grid, "8 RK4 + 8 AB4 steps, tri 1 vs 0: max|d rho|", np.abs(states[1] - states[0]).max())
    assert np.array_equal(states[1], states[0])


@pytest.mark.parametrize("name,grid", [("ipa14_hex120", [10, 9]), ("ipa6_hex96", [16, 16])])
def test_dft_fields_agree_and_meet_oracle(model_dir, name, grid):
    p = _model(model_dir, name, grid)
    o = oracle_py.Oracle(p)
    H0 = o.get("H0k")
    dev = np.abs(H0 - H0.conj().transpose(0, 2, 1)).max() / np.abs(H0).max()
    _say(name, grid, "H0(k) Hermiticity defect / max|H0| =", dev, "=", dev / np.finfo(float).eps, "ulp")
    eng = _engine(o)
    assert eng.get_tuning().tri_fields == 1
    rho = o.rho + 0.1 * _rho(o.Nk, o.nb, 5)
    E = [3e-3, 2e-3, 0.0]
    f_ref = o.rhs_E(E, rho)
    scale = np.abs(f_ref).max()
    f = {}
    for on in (1, 0):
        _set_tri(eng, on)
        f[on] = eng.rhs(E, rho)
        err = np.abs(f[on] - f_ref).max()
        _say(name, grid, "tri", on, "max|df| vs oracle / max|f|", err / scale)
        assert err <= 1e-12 * scale
    d = np.abs(f[1] - f[0]).max() / scale
    _say(name, grid, "tri 1 vs 0: max|df| / max|f| =", d)
    assert d <= 64 * np.finfo(float).eps


def test_not_applicable_contexts(model_dir):
    p = _model(model_dir, "ipa14_hex120", [10, 9])
    o = oracle_py.Oracle(p)
    cases = {"hermitian layout": _engine(o, layout=HERM)}
    rng = np.random.default_rng(1)
    H0 = o.get("H0k")
    bad = H0 + 1e-9 * np.abs(H0).max() * rng.standard_normal(H0.shape)  # anti-Hermitian part of 1e-9
    cases["non-Hermitian H0"] = _engine(o, H0k=bad)
    q = _model(model_dir, "hsex8_hex90", [9, 11])
    oq = oracle_py.Oracle(q)
    cases["exchange"] = _engine(oq, coulomb=True)
    for what, eng in cases.items():
        t = eng.get_tuning()
        _say(what, "tri_fields", t.tri_fields)
        assert t.tri_fields == 0
        t.tri_fields = 1
        with pytest.raises(edus.EdusError, match="tri_fields"):
            eng.set_tuning(t)


def test_hbn_simulation_storage():
    """The hBN fixture through the Simulation set-up: the storage chosen (its
    H(R) file may hold H(-R) = H(R)^dagger only to its printed digits) and
    parity of a few RK4 steps with the oracle either way."""
    p = hbn_params(grid=[12, 12, 1])
    p["lasers"][0]["t0"] = -0.5
    sim = edus.Simulation(p)
    tri = sim.engine.get_tuning().tri_fields
    o = oracle_py.Oracle(p)
    sim.propagate(5)
    o.propagate(5)
    d = np.abs(sim.engine.get_state() - o.rho).max()
    _say("hBN 12x12: tri_fields", tri, "5 RK4 steps max|d rho| vs oracle", d)
    assert d <= 1e-12
