"""GPU parity at the benchmark shapes themselves, against the oracle: the
14-band 120 x 120 RK4 workload and the 8-band 90 x 90 AB4 workload with
screened exchange, each set up by the library's host code from the same
tight-binding files as the oracle's.

Bar: one right-hand side within 1e-11 relative; a short propagation with
max|d rho| <= 1e-10 and observables within 1e-10 (current 1e-9 relative).
"""
import os
import sys

import numpy as np
import pytest

import edus
from edus import models

oracle_py = pytest.importorskip("oracle_py")

pytestmark = pytest.mark.gpu

FULL, HERM = edus.LAYOUT_FULL, edus.LAYOUT_HERMITIAN


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    p = models.workload("ipa14_hex120", str(tmp_path_factory.mktemp("c1")))
    _say("c1 oracle set-up")
    return p, oracle_py.Oracle(p)


def _rhs_check(o, sim, layout, E, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((o.Nk, o.nb, o.nb)) + 1j * rng.standard_normal((o.Nk, o.nb, o.nb))
    rho = o.rho + 0.05 * 0.5 * (a + a.conj().transpose(0, 2, 1))
    f_ref = o.rhs_E(E, rho)
    f = sim.engine.rhs(E, rho)
    if layout == HERM:
        iu = np.triu_indices(o.nb)
        f, f_ref = f[:, iu[0], iu[1]], f_ref[:, iu[0], iu[1]]
    err, scale = np.abs(f - f_ref).max(), max(1.0, np.abs(f_ref).max())
    _say("  E", E, "max|df|", err, "scale", scale)
    assert err <= 1e-11 * scale


@pytest.mark.parametrize("layout", [FULL, HERM])
def test_c1_rhs_at_benchmark_shape(c1, layout):
    p, o = c1
    sim = edus.Simulation(p, layout=layout)
    for seed, E in ((1, [0.02, 0.0, 0.0]), (2, [3e-3, -4e-3, 0.0])):
        _rhs_check(o, sim, layout, E, seed)


def _compare(ref, got, drho):
    for key in ("time", "laser", "pop_bloch", "pop_wannier", "velocity"):
        err = np.abs(got[key] - ref[key]).max()
        _say("  ", key, "max abs err", err)
        assert err <= 1e-10, key
    vref = np.abs(ref["velocity"]).max()
    if vref > 0:
        assert np.abs(got["velocity"] - ref["velocity"]).max() <= 1e-9 * vref + 1e-14
    _say("   max|d rho|", drho)
    assert drho <= 1e-10


def test_c1_rk4_steps_at_benchmark_shape(c1):
    """The benchmark's own step (FULL, FP64) for 12 steps, every step
    printed, against the oracle's RK4; the pulse starts 10 fs early so the
    steps run at its peak field."""
    p = dict(c1[0], printresolution=1, printresolution_pulse=1)
    p["lasers"] = [dict(p["lasers"][0], t0=-10.0)]
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    ref = o.propagate(12)
    got = sim.propagate(12)
    _compare(ref, got, np.abs(sim.engine.get_state() - o.rho).max())
    tr, herm = sim.engine.check()
    _say("  monitors", tr, herm)
    assert tr < 1e-11


def test_c2_ab4_exchange_steps_at_benchmark_shape(tmp_path):
    """The 8-band, 90 x 90 AB4 workload with Hartree + screened exchange, 16
    steps (the zero-history start included) against the oracle, at the peak
    of its two pulses."""
    p = models.workload("hsex8_hex90", str(tmp_path))
    p["printresolution"] = p["printresolution_pulse"] = 1
    p["lasers"] = [dict(las, t0=-7.5) for las in p["lasers"]]  # at the pulses' peak
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p)
    ref = o.propagate(16)
    got = sim.propagate(16)
    _compare(ref, got, np.abs(sim.engine.get_state() - o.rho).max())


def test_c1_mixed_precision_at_benchmark_shape(c1):
    """Mixed precision at the 14-band 120 x 120 shape (H0 and xi stored in
    FP32, FP32-MFMA commutator and rotations; gradient, state and updates
    FP64) against the FP64 oracle: one right-hand side, then 12 RK4 steps
    at the pulse peak with populations and current.  Bars: RHS 2e-6 of
    max|f|; rho 1e-6; populations 1e-6; current 1e-4 relative (measured
    values printed)."""
    p = dict(c1[0], printresolution=1, printresolution_pulse=1)
    p["lasers"] = [dict(p["lasers"][0], t0=-10.0)]
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, precision=edus.PREC_MIXED)
    assert sim.engine.get_tuning().comm_kind == 2
    rng = np.random.default_rng(3)
    a = rng.standard_normal((o.Nk, o.nb, o.nb)) + 1j * rng.standard_normal((o.Nk, o.nb, o.nb))
    rho = o.rho + 0.05 * 0.5 * (a + a.conj().transpose(0, 2, 1))
    E = [0.02, 0.0, 0.0]
    f_ref = o.rhs_E(E, rho)
    rel = np.abs(sim.engine.rhs(E, rho) - f_ref).max() / np.abs(f_ref).max()
    _say("c1 mixed: RHS max|df| / max|f| =", rel)
    assert 1e-12 < rel <= 2e-6
    ref = o.propagate(12)
    got = sim.propagate(12)
    drho = np.abs(sim.engine.get_state() - o.rho).max()
    dpop = np.abs(got["pop_bloch"] - ref["pop_bloch"]).max()
    dvel = np.abs(got["velocity"] - ref["velocity"]).max() / np.abs(ref["velocity"]).max()
    _say("c1 mixed, 12 RK4 steps: max|d rho|", drho, "max|d pop|", dpop, "rel d current", dvel)
    assert drho <= 1e-6 and dpop <= 1e-6 and dvel <= 1e-4


def test_c2_mixed_precision_at_benchmark_shape(tmp_path):
    """Mixed precision at the 8-band 90 x 90 AB4 shape with Hartree + screened
    exchange (H0, which holds the folded ground-state self energy, and the
    exchange self energy stay FP64; the products run on the FP32 matrix cores)
    against the FP64 oracle, 16 steps at the pulses' peak.  Bars: rho 1e-6;
    populations 1e-6; current 1e-4 relative (measured values printed)."""
    p = models.workload("hsex8_hex90", str(tmp_path))
    p["printresolution"] = p["printresolution_pulse"] = 1
    p["lasers"] = [dict(las, t0=-7.5) for las in p["lasers"]]
    o = oracle_py.Oracle(p)
    sim = edus.Simulation(p, precision=edus.PREC_MIXED)
    ref = o.propagate(16)
    got = sim.propagate(16)
    drho = np.abs(sim.engine.get_state() - o.rho).max()
    dpop = np.abs(got["pop_bloch"] - ref["pop_bloch"]).max()
    dvel = np.abs(got["velocity"] - ref["velocity"]).max() / np.abs(ref["velocity"]).max()
    _say("c2 mixed, 16 AB4 steps: max|d rho|", drho, "max|d pop|", dpop, "rel d current", dvel)
    assert drho <= 1e-6 and dpop <= 1e-6 and dvel <= 1e-4
