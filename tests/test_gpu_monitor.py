"""GPU: the invariant monitor fused into the last stage kernel reports the
trace and Hermiticity defects of the state it wrote, exactly.

A step with dt = 0 leaves the state unchanged (RK4: every stage adds 0 * f;
AB: y + 0 * f + 0 * history), so after one such step edus_check must return
    trace_dev = max_k |Tr rho(k) - nf|,
    herm_dev  = max_k max_{n,m} |rho_nm(k) - conj(rho_mn(k))|
of the planted state, computed here from the state read back (in the
HERMITIAN layout that is the stored upper triangle, completed).  A NaN
anywhere must trip both monitors (+inf).
"""
import sys

import numpy as np
import pytest

import edus

pytestmark = pytest.mark.gpu

FULL, HERM = edus.LAYOUT_FULL, edus.LAYOUT_HERMITIAN


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(nb, n0, n1, layout, precision, nf, seed):
    rng = np.random.default_rng(seed)
    h = rng.standard_normal((n0 * n1, nb, nb)) + 1j * rng.standard_normal((n0 * n1, nb, nb))
    h = 0.5 * (h + h.conj().transpose(0, 2, 1))
    lat = np.array([[1.0, 0.0, 0.0], [0.5, 0.8, 0.0], [0.0, 0.0, 1.0]])
    return edus.Engine(nb, n0, n1, H0k=h, A=lat, filledbands=nf, layout=layout, precision=precision)


def _planted(nk, nb, nf, seed):
    """A state near a Hermitian one of trace nf, with defects of known size at a few k."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))
    rho = 0.05 * (a + a.conj().transpose(0, 2, 1))
    rho[:, np.arange(nb), np.arange(nb)] = rho[:, np.arange(nb), np.arange(nb)].real + nf / nb
    rho += 1e-9 * (rng.standard_normal(rho.shape) + 1j * rng.standard_normal(rho.shape))
    rho[nk // 3, 0, nb - 1] += 3e-4 - 2e-4j          # an off-diagonal defect
    rho[nk - 1, nb // 2, nb // 2] += 1e-3 + 5e-5j     # trace and an imaginary diagonal
    return rho


def _expected(rho, nf):
    tr = np.abs(np.einsum("kii->k", rho) - nf).max()
    herm = np.abs(rho - rho.conj().transpose(0, 2, 1)).max()
    return tr, herm


@pytest.mark.parametrize("nb,grid", [(14, (9, 10)), (5, (16, 12)), (1, (7, 6))])
@pytest.mark.parametrize("layout", [FULL, HERM])
@pytest.mark.parametrize("precision", [edus.PREC_FP64, edus.PREC_MIXED])
@pytest.mark.parametrize("solver", ["RK4", "AB4"])
def test_monitor_reports_planted_defects(nb, grid, layout, precision, solver):
    nf = max(1, nb // 2)
    _say("case", nb, grid, layout, precision, solver)
    eng = _engine(nb, grid[0], grid[1], layout, precision, nf, seed=nb)
    eng.set_state(_planted(eng.Nk, nb, nf, seed=7 + nb))
    before = eng.get_state()
    eng.check()  # clear
    if solver == "RK4":
        eng.rk4_steps(0.0, np.zeros((1, 3, 3)))
    else:
        eng.ab_steps(0.0, 4, np.zeros((1, 3)))
    after = eng.get_state()
    assert np.array_equal(after, before)  # dt = 0 leaves the state as it was
    tr, herm = eng.check()
    etr, eherm = _expected(after, nf)
    _say("  monitor", tr, herm, "expected", etr, eherm)
    assert abs(tr - etr) <= 1e-12 * max(etr, 1.0)
    assert abs(herm - eherm) <= 1e-12 * max(eherm, 1e-300)
    assert eng.check() == (0.0, 0.0)  # reading clears it


@pytest.mark.parametrize("layout", [FULL, HERM])
def test_monitor_trips_on_nan(layout):
    nb, nf = 6, 3
    eng = _engine(nb, 8, 9, layout, edus.PREC_FP64, nf, seed=3)
    rho = _planted(eng.Nk, nb, nf, seed=4)
    rho[11, 2, 3] = np.nan
    eng.set_state(rho)
    eng.check()
    eng.rk4_steps(0.0, np.zeros((1, 3, 3)))
    tr, herm = eng.check()
    _say("nan case", layout, tr, herm)
    assert np.isinf(tr) and np.isinf(herm)


@pytest.mark.parametrize("solver", ["RK4", "AB4"])
def test_monitor_limits_stop_the_step_calls(solver):
    """edus_monitor_limits: a step call whose states exceed a set limit
    returns EDUS_ERR_MONITOR (5) with the values in the message; within the
    limits it succeeds; a NaN fails whatever the limits; limits of 0 check
    nothing."""
    nb, nf = 5, 2
    eng = _engine(nb, 16, 12, FULL, edus.PREC_FP64, nf, seed=5)
    rho = _planted(eng.Nk, nb, nf, seed=6)
    etr, eherm = _expected(rho, nf)

    def step():
        if solver == "RK4":
            eng.rk4_steps(0.0, np.zeros((1, 3, 3)))
        else:
            eng.ab_steps(0.0, 4, np.zeros((1, 3)))

    eng.set_state(rho)
    eng.check()
    eng.monitor_limits(2.0 * etr, 2.0 * eherm)
    step()  # within the limits
    eng.check()
    eng.monitor_limits(0.5 * etr, 0.0)  # the trace defect is twice this limit
    with pytest.raises(edus.EdusError, match="trace monitor") as e:
        step()
    assert e.value.code == 5
    eng.check()
    eng.monitor_limits(0.0, 0.5 * eherm)
    with pytest.raises(edus.EdusError, match="Hermiticity monitor"):
        step()
    eng.check()
    eng.monitor_limits(0.0, 0.0)  # off
    step()
    eng.check()
    bad = rho.copy()
    bad[7, 1, 2] = np.nan
    eng.set_state(bad)
    eng.monitor_limits(1e300, 1e300)
    with pytest.raises(edus.EdusError) as e:
        step()
    assert e.value.code == 5
    with pytest.raises(edus.EdusError):
        eng.monitor_limits(-1.0, 0.0)
