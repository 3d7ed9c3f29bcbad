"""Wannier <-> Bloch gauge rotation on the matrix cores (edus_rotate,
Operator::go_to_bloch / go_to_wannier, src/Operator/Operator.hpp:528-564)
against the FP64 oracle's rotation, which tests/test_oracle.py pins to numpy
and to the eigensystem's closed forms.

Bar: FP64 within 1e-12 of max|O|; the FP32 matrix-core rotation of a
mixed-precision context within 2e-6 of max|O| (measured error printed).
Covered: every band count class (1, 2, 3, 5, 8, 13, 14, 16), odd and even
grids, both storage layouts, both directions, the round trip, the state in
the Bloch gauge and the device-side initial condition."""
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
    tb = os.path.join(tmp, f"rot{nb}")
    models.write_tb(tb + "_tb.dat", m)
    return config.compose({"tb_file": tb, "grid": [grid[0], grid[1], 1], "filledbands": max(1, nb // 2),
                           "dt": 0.05, "dt_units": "autime", "finaltime": 5.0, "finaltime_units": "fs",
                           "lasers": [dict(models.laser(0.01, 1.4, 5.0, (1, 0.5, 0)), t0=-2.5)]})


def _engine(o, p, layout, precision=edus.PREC_FP64):
    return edus.Engine(o.nb, o.N0, o.N1, H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")),
                       A=o.lattice(), filledbands=p["filledbands"], U=o.get("U"), layout=layout,
                       precision=precision)


def _operator(nk, nb, seed, hermitian):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))
    return 0.5 * (a + a.conj().transpose(0, 2, 1)) if hermitian else a


CASES = [(1, [6, 5], "square"), (2, [9, 8], "hex"), (3, [2, 2], "hex"), (5, [7, 4], "square"),
         (8, [10, 9], "hex"), (13, [4, 7], "square"), (14, [12, 10], "hex"), (16, [11, 6], "square")]


@pytest.mark.parametrize("layout", [FULL, HERM])
@pytest.mark.parametrize("nb,grid,kind", CASES)
def test_rotation_matches_oracle(tmp_path, nb, grid, kind, layout):
    p = _params(str(tmp_path), nb, grid, kind)
    _say("case nb", nb, "grid", grid, "layout", layout)
    o = oracle_py.Oracle(p)
    eng = _engine(o, p, layout)
    x = _operator(o.Nk, nb, nb, hermitian=(layout == HERM))
    for to_bloch in (True, False):
        ref = o.rotate(to_bloch, x)
        got = eng.rotate(to_bloch, x)
        err = np.abs(got - ref).max()
        _say("  to_bloch", to_bloch, "max|dO|", err, "scale", np.abs(ref).max())
        assert err <= 1e-12 * max(1.0, np.abs(x).max())
    back = eng.rotate(False, eng.rotate(True, x))
    assert np.abs(back - x).max() <= 1e-12 * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("nb,grid,kind", [(2, [9, 8], "hex"), (8, [10, 9], "hex"), (14, [12, 10], "hex"),
                                          (16, [11, 6], "square")])
def test_rotation_fp32_error(tmp_path, nb, grid, kind):
    """Mixed precision: FP32 operands and accumulation on the matrix cores."""
    p = _params(str(tmp_path), nb, grid, kind)
    o = oracle_py.Oracle(p)
    eng = _engine(o, p, FULL, edus.PREC_MIXED)
    x = _operator(o.Nk, nb, 7, hermitian=True)
    for to_bloch in (True, False):
        ref = o.rotate(to_bloch, x)
        err = np.abs(eng.rotate(to_bloch, x) - ref).max() / np.abs(x).max()
        _say("nb", nb, "to_bloch", to_bloch, "FP32 rotation error (of max|O|)", err)
        assert 1e-12 < err <= 2e-6  # FP32, not silently FP64


def test_state_in_bloch_gauge_and_initial_condition(tmp_path):
    """edus_get_state_bloch of a propagated state against the oracle's
    rotation of its own state; the simulation's device-side initial
    condition against the oracle's (U diag(filled) U^dagger)."""
    p = _params(str(tmp_path), 6, [9, 10], "hex")
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    assert np.abs(sim.engine.get_state() - o.rho).max() < 1e-13
    assert np.abs(sim.get("rho0") - o.rho).max() < 1e-13
    sim.propagate(20)
    o.propagate(20)
    assert np.abs(sim.engine.get_state() - o.rho).max() < 1e-10
    rb = sim.engine.get_state_bloch()
    # the off-diagonal of U^dagger rho U depends on the phases of U's columns,
    # which the library's and the oracle's eigensolvers choose independently:
    # compare against the oracle's state rotated with the library's own U
    # (numpy), and the gauge-invariant diagonal against the oracle's rotation
    Us = sim.get("U")
    ref = Us.conj().transpose(0, 2, 1) @ o.rho @ Us
    err = np.abs(rb - ref).max()
    derr = np.abs(np.einsum("knn->kn", rb) - np.einsum("knn->kn", o.rotate(True, o.rho))).max()
    _say("DMk_bloch max|d|", err, "diagonal vs the oracle's rotation", derr)
    assert err < 1e-10 and derr < 1e-10


def test_state_bloch_into_device_buffer(tmp_path):
    """edus_get_state_bloch_device writes the same [Nk][nb][nb] Bloch-gauge
    state as edus_get_state_bloch, into a device buffer on the context's
    stream."""
    from edus import hip_rt

    p = models.workload("ipa6_hex96", str(tmp_path), grid=[12, 10, 1])
    sim = edus.Simulation(p)
    eng = sim.engine
    sim.steps(3)
    ref = eng.get_state_bloch()
    buf = hip_rt.DeviceBuffer(ref.nbytes)
    eng.get_state_bloch_device(buf.address)
    eng.synchronize()
    got = buf.to_array(np.complex128, ref.shape)
    assert np.array_equal(got, ref)
