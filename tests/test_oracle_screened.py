"""Pins the oracle's screened-potential application (src/Screened_potential/
Screened_potential.cpp) before it is used as the GPU checker.

  - the Lindhard sum (ResponseFunction, :17-58) against hand-derived closed
    forms of gauge-phased two-band blocks (screened_models.py), which fix the
    k + q direction, the order of the eigenvector factors and the denominator;
  - invariants: X(q) Hermitian for every q, and unchanged by any phase of the
    eigenvector columns (the response cannot depend on the gauge of U);
  - the R-space steps on the reference's hBN input: X(R) against numpy's FFT
    of X(k), and V(R) = 1/|r_a + R - r_b| against the distances of the
    reference's own DM0.txt (ci-test/outputs/hBN), then the elementwise
    Eps / InvEps / W relations of Calculate (:139-199).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, hbn_params
import screened_models as sm

oracle_py = pytest.importorskip("oracle_py")


@pytest.mark.parametrize("nblocks,N0,N1", [(1, 6, 5), (1, 7, 4), (3, 5, 4), (2, 4, 6)])
def test_lindhard_closed_form(nblocks, N0, N1):
    U, eps, nf = sm.gauge_block_model(nblocks, N0, N1)
    X = oracle_py.lindhard(U, eps, N0, N1, nf)
    ref = sm.closed_form(nblocks, N0, N1)
    scale = np.abs(ref).max()
    assert scale > 1e-3
    assert np.abs(X - ref).max() <= 1e-13 * scale


def test_closed_form_distinguishes_the_shift_direction():
    """The closed form itself is sensitive to k + q versus k - q: X(q) and
    X(-q) differ, so a mirrored shift in the oracle would fail the test above."""
    N0, N1 = 6, 5
    ref = sm.closed_form(1, N0, N1)
    q = 1 * N1 + 2
    mq = ((N0 - 1) % N0) * N1 + (N1 - 2) % N1
    assert np.abs(ref[q] - ref[mq]).max() > 1e-3 * np.abs(ref).max()


@pytest.mark.parametrize("nb,nf,N0,N1", [(3, 1, 4, 5), (4, 2, 5, 4), (5, 3, 3, 4)])
def test_lindhard_hermitian_and_gauge_invariant(nb, nf, N0, N1):
    U, w = sm.random_bands(nb, N0, N1, seed=nb)
    X = oracle_py.lindhard(U, w, N0, N1, nf)
    scale = np.abs(X).max()
    assert np.abs(X - X.conj().transpose(0, 2, 1)).max() <= 1e-13 * scale
    rng = np.random.default_rng(7)
    Ug = U * np.exp(1j * rng.uniform(0, 2 * np.pi, size=(N0 * N1, 1, nb)))
    Xg = oracle_py.lindhard(Ug, w, N0, N1, nf)
    assert np.abs(Xg - X).max() <= 1e-13 * scale


def test_lindhard_no_valence_or_conduction_is_zero():
    U, w = sm.random_bands(3, 3, 3, seed=1)
    assert np.abs(oracle_py.lindhard(U, w, 3, 3, 0)).max() == 0.0
    assert np.abs(oracle_py.lindhard(U, w, 3, 3, 3)).max() == 0.0


@pytest.fixture(scope="module")
def hbn_screened():
    o = oracle_py.Oracle(hbn_params())
    return o, o.screened_potential()


def test_hbn_response_to_R_is_the_fft(hbn_screened):
    o, f = hbn_screened
    Xk = oracle_py.lindhard(o.get("U"), o.eigenvalues(), o.N0, o.N1, 1)
    grid = Xk.reshape(o.N0, o.N1, o.nb, o.nb)
    XR = np.fft.fft2(grid, axes=(0, 1)).reshape(o.Nk, o.nb, o.nb) / o.Nk  # sign -1, 1/N
    assert np.abs(f["X"] - XR).max() <= 1e-13 * np.abs(XR).max()
    assert np.abs(XR).max() > 1e-4


def test_hbn_bare_coulomb_matches_reference_distances(hbn_screened):
    """DM0.txt's first column is |r_a + R - r_b| (bohr) on the same grid and
    order, written by the reference (10 significant digits)."""
    _, f = hbn_screened
    d = np.loadtxt(os.path.join(GOLDEN, "hBN", "DM0.txt"))[:, 0]
    V = f["V"].reshape(-1)
    far = d > 1e-5
    assert np.all(V[~far] == 0)
    assert np.abs(V[far].real * d[far] - 1.0).max() < 1e-8
    assert np.abs(V.imag).max() == 0


def test_hbn_permittivity_inverse_and_screened_interaction(hbn_screened):
    _, f = hbn_screened
    X, V, Eps, Inv, W = (f[k] for k in ("X", "V", "Eps", "InvEps", "W"))
    assert np.abs(Eps - (1.0 - V * X)).max() <= 1e-15 * max(1.0, np.abs(Eps).max())
    small = np.abs(Eps) < 1e-3
    assert np.all(Inv[small] == 0)
    assert np.abs(Inv[~small] * Eps[~small] - 1.0).max() < 1e-13
    assert np.abs(W - Inv * V).max() <= 1e-15 * max(1.0, np.abs(W).max())
    # on-site diagonal: V = 0 there, so Eps = 1 and W = 0
    assert np.abs(Eps[0].diagonal() - 1.0).max() == 0 and np.abs(W[0].diagonal()).max() == 0
