"""CPU: the oracle's Peierls-gauge branch pinned before it judges the GPU.

- With E = 0 and A.a_j = 2 pi m_j / N_j the phase exp(i A.R) on H(R) is a shift
  of the k grid, so the right-hand side is -i[H0(k + shift), rho]: pins the
  sign and the Gamma-centred R convention (a flipped sign shifts the other way).
- The vector potential against -int E dt by quadrature (Simpson over the
  reference's call sequence t, t+dt/2, t+dt).
- Two lasers: the reference's shared integration state (every laser returns
  the first laser's integral).
"""
import numpy as np
from scipy import integrate

import oracle_py
from conftest import hbn_params


def _oracle(**kw):
    p = hbn_params(grid=[12, 10, 1], peierls=True, **kw)
    return oracle_py.Oracle(p)


def test_peierls_phase_is_a_grid_shift():
    o = _oracle()
    A = o.lattice()  # columns: lattice vectors (bohr)
    m = (2, -3)
    # A.a_j = 2 pi m_j / N_j  ->  solve for the Cartesian A (in-plane)
    M2 = A[:2, :2].T
    theta = np.array([2 * np.pi * m[0] / o.N0, 2 * np.pi * m[1] / o.N1])
    Av = np.zeros(3)
    Av[:2] = np.linalg.solve(M2, theta)
    rng = np.random.default_rng(0)
    rho = rng.standard_normal((o.Nk, 2, 2)) + 1j * rng.standard_normal((o.Nk, 2, 2))
    f = o.rhs_EA([0.0, 0.0, 0.0], Av, rho)
    H = o.get("H0k").reshape(o.N0, o.N1, 2, 2)
    # H(R) exp(iA.R) -> sum_R H(R) e^{i(k+A).R}: k index shifted by +m
    Hs = np.roll(H, shift=(-m[0], -m[1]), axis=(0, 1)).reshape(o.Nk, 2, 2)
    ref = -1j * (Hs @ rho - rho @ Hs)
    assert np.abs(f - ref).max() <= 1e-11 * np.abs(ref).max()
    # the opposite shift is far away (the check can fail)
    Hw = np.roll(H, shift=(m[0], m[1]), axis=(0, 1)).reshape(o.Nk, 2, 2)
    assert np.abs(f - (-1j * (Hw @ rho - rho @ Hw))).max() > 1e-3


def test_vector_potential_is_minus_the_field_integral():
    """Simpson over each half step: error against quadrature O(dt^4)."""
    T = 24.0

    def err(dt):
        o = _oracle()
        t, worst, amax = 0.0, 0.0, 0.0
        while t < T - 1e-12:
            for frac in (0.0, 0.5, 0.5, 1.0):  # RK4 stage times, as SourceTerm calls it
                Av = o.vector_potential(t + frac * dt)
            t += dt
            ref = np.array([-integrate.quad(lambda x: o.E(x)[d], 0.0, t, limit=400, epsabs=1e-15)[0]
                            for d in range(2)])
            worst = max(worst, np.abs(Av[:2] - ref).max())
            amax = max(amax, np.abs(ref).max())
        return worst / amax

    e1, e2 = err(0.4), err(0.2)
    assert e2 <= 3e-6
    assert 12.0 <= e1 / e2 <= 20.0, (e1, e2)


def test_vector_potential_shared_state_over_lasers():
    base = hbn_params()
    las2 = dict(base["lasers"][0])
    las2["polarization"] = [0, 1, 0]
    o2 = oracle_py.Oracle(hbn_params(grid=[6, 6, 1], peierls=True, lasers=[base["lasers"][0], las2]))
    o1 = oracle_py.Oracle(hbn_params(grid=[6, 6, 1], peierls=True, lasers=[base["lasers"][0]]))
    for tt in np.linspace(0.0, 20.0, 41):
        a2 = o2.vector_potential(tt)
        a1 = o1.vector_potential(tt)
        assert np.allclose(a2, 2.0 * a1, rtol=0, atol=1e-15)


def test_laser_A_is_recorded_at_every_print():
    """Laser_A.txt: the vector potential is evaluated at every print step in
    either gauge (Simulation.cpp:425); in the length gauge those are its only
    calls, so it is Simpson's rule over the print intervals."""
    p = hbn_params(grid=[6, 6, 1])
    o = oracle_py.Oracle(p)
    res = o.propagate(40)
    t, la = res["time"], res["laser_A"]
    assert la.shape == (len(t), 3)
    ref = np.zeros(3)
    prev = 0.0
    for i, ti in enumerate(t):
        dT = ti - prev
        if dT >= 1e-7:
            ref = ref - dT / 6.0 * (o.E(prev) + 4 * o.E(prev + dT / 2) + o.E(ti))
            prev = ti
        assert np.abs(la[i] - ref).max() <= 1e-15 + 1e-12 * np.abs(ref).max()
