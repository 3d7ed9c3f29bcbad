"""Edge shapes of the engine against the oracle: every band count the
kernels are instantiated for at its ends and in between (nb = 1, odd counts,
the maximum 16), the smallest grid (2 x 2), prime and Stockham-only axis
lengths (7, 11, 13, 26), both layouts.

Bar: one right-hand side within 1e-11 relative; RK4 propagations with
max|d rho| <= 1e-10 and observables within 1e-10."""
import os
import sys

import numpy as np
import pytest

import edus
from edus import config, models

oracle_py = pytest.importorskip("oracle_py")

pytestmark = pytest.mark.gpu

FULL, HERM = edus.LAYOUT_FULL, edus.LAYOUT_HERMITIAN


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(tmp, nb, grid, kind="square", seed=3):
    ons = list(np.linspace(-2.0, 2.0, nb)) if nb > 1 else [0.5]
    m = models.random_model(nb, kind, 2.0, 0.6, ons, 0.08, 0.03, seed)
    tb = os.path.join(tmp, f"edge{nb}")
    models.write_tb(tb + "_tb.dat", m)
    return config.compose({"tb_file": tb, "grid": [grid[0], grid[1], 1], "filledbands": max(1, nb // 2),
                           "dt": 0.05, "dt_units": "autime", "finaltime": 5.0, "finaltime_units": "fs",
                           "lasers": [dict(models.laser(0.01, 1.4, 5.0, (1, 0.5, 0)), t0=-2.5)]})


def _engine(o, p, layout):
    return edus.Engine(o.nb, o.N0, o.N1, H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")),
                       A=o.lattice(), filledbands=p["filledbands"], U=o.get("U"), layout=layout)


CASES = [(1, [6, 5], "square"), (3, [2, 2], "hex"), (5, [7, 4], "square"), (11, [5, 6], "hex"),
         (13, [4, 7], "square"), (15, [3, 5], "hex"), (16, [11, 26], "square")]


@pytest.mark.parametrize("layout", [FULL, HERM])
@pytest.mark.parametrize("nb,grid,kind", CASES)
def test_rhs_edge_shapes(tmp_path, nb, grid, kind, layout):
    p = _params(str(tmp_path), nb, grid, kind)
    _say("case nb", nb, "grid", grid, "layout", layout)
    o = oracle_py.Oracle(p)
    eng = _engine(o, p, layout)
    rng = np.random.default_rng(nb)
    for E in ([4e-3, -3e-3, 0.0], [0.0, 5e-3, 0.0]):
        a = rng.standard_normal((o.Nk, nb, nb)) + 1j * rng.standard_normal((o.Nk, nb, nb))
        rho = 0.5 * (a + a.conj().transpose(0, 2, 1))
        f_ref = o.rhs_E(E, rho)
        # every FP64 commutator engine and tile at this shape
        for kind, tile in ((0, 0), (1, 0), (1, 1), (1, 2)):
            t = eng.get_tuning()
            t.comm_kind, t.comm_tile, t.comm_tile_ab = kind, tile, tile
            eng.set_tuning(t)
            f = eng.rhs(E, rho)
            fr = f_ref
            if layout == HERM:
                iu = np.triu_indices(nb)
                f, fr = f[:, iu[0], iu[1]], f_ref[:, iu[0], iu[1]]
            err, scale = np.abs(f - fr).max(), max(1.0, np.abs(fr).max())
            _say("  E", E, "engine", kind, "tile", tile, "max|df|", err, "scale", scale)
            assert err <= 1e-11 * scale


@pytest.mark.parametrize("nb,grid,kind", [c for c in CASES if c[0] != 3] + [(3, [3, 4], "hex")])
def test_rk4_edge_shapes(tmp_path, nb, grid, kind):
    p = _params(str(tmp_path), nb, grid, kind)
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    ref = o.propagate(12)
    got = sim.propagate(12)
    for key in ("time", "laser", "pop_bloch", "pop_wannier", "velocity"):
        err = np.abs(got[key] - ref[key]).max()
        assert err <= 1e-10, key
    drho = np.abs(sim.engine.get_state() - o.rho).max()
    _say("nb", nb, "grid", grid, "max|d rho|", drho)
    assert drho <= 1e-10
