"""GPU: edus_fourier (Operator::go_to_R / go_to_k, Operator.hpp:566-606,
through FourierTransform::fft, fftPair.cpp:89-107) and the C++ seam classes.

- k -> R and R -> k of random operator fields against numpy's FFT (sign -1
  scaled by 1/Nk, sign +1 unscaled), odd, even and mixed-radix grids, and the
  round trip;
- H0(k), built by the set-up's explicit DFT of the tight-binding H(R), goes
  back to H(R) at the folded R positions and is zero elsewhere: a flipped
  sign would put H(R)^dagger there, a missing 1/Nk would scale it;
- the arguments the entry point refuses.
"""
import os
import sys

import numpy as np
import pytest

import edus
from edus import config, models

oracle_py = pytest.importorskip("oracle_py")

pytestmark = pytest.mark.gpu

HARTREE_EV = 2 * 2.1798723611035e-18 / 1.602176634e-19


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup(tmp, nb, grid, kind, seed=2, layout=edus.LAYOUT_FULL):
    ons = list(np.linspace(-2.0, 2.0, nb)) if nb > 1 else [0.5]
    m = models.random_model(nb, kind, 2.0, 0.6, ons, 0.08, 0.03, seed)
    tb = os.path.join(tmp, f"ft{nb}")
    models.write_tb(tb + "_tb.dat", m)
    p = config.compose({"tb_file": tb, "grid": [grid[0], grid[1], 1], "filledbands": max(1, nb // 2),
                        "dt": 0.05, "dt_units": "autime", "finaltime": 5.0, "finaltime_units": "fs",
                        "lasers": [models.laser(0.01, 1.4, 5.0, (1, 0, 0))]})
    o = oracle_py.Oracle(p)
    eng = edus.Engine(o.nb, o.N0, o.N1, H0k=o.get("H0k"), xik=(o.get("rk0"), o.get("rk1"), o.get("rk2")),
                      A=o.lattice(), filledbands=p["filledbands"], U=o.get("U"), layout=layout)
    return m, o, eng


@pytest.mark.parametrize("nb,grid,kind", [(2, [12, 15], "square"), (3, [10, 9], "hex"), (6, [16, 16], "hex"),
                                          (14, [120, 120], "hex")])
def test_fourier_matches_numpy(tmp_path, nb, grid, kind):
    _, o, eng = _setup(str(tmp_path), nb, grid, kind)
    assert eng.info() == (nb, grid[0], grid[1], edus.LAYOUT_FULL)
    rng = np.random.default_rng(7)
    N0, N1 = grid
    x = rng.standard_normal((N0 * N1, nb, nb)) + 1j * rng.standard_normal((N0 * N1, nb, nb))
    ref_R = np.fft.fft2(x.reshape(N0, N1, nb, nb), axes=(0, 1)).reshape(x.shape) / (N0 * N1)
    got_R = eng.fourier(-1, x)
    eR = np.abs(got_R - ref_R).max() / np.abs(ref_R).max()
    ref_k = np.fft.ifft2(x.reshape(N0, N1, nb, nb), axes=(0, 1)).reshape(x.shape) * (N0 * N1)
    got_k = eng.fourier(+1, x)
    ek = np.abs(got_k - ref_k).max() / np.abs(ref_k).max()
    back = eng.fourier(+1, got_R)
    eb = np.abs(back - x).max() / np.abs(x).max()
    _say(nb, grid, "k->R vs numpy", eR, "R->k vs numpy", ek, "round trip", eb)
    assert eR <= 1e-14 and ek <= 1e-14 and eb <= 1e-14
    # the state is left alone
    before = eng.get_state()
    eng.fourier(-1, x)
    assert np.array_equal(eng.get_state(), before)


@pytest.mark.parametrize("nb,grid,kind", [(2, [12, 15], "square"), (6, [16, 16], "hex"), (5, [9, 10], "hex")])
def test_h0k_goes_back_to_the_tight_binding_h(tmp_path, nb, grid, kind):
    m, o, eng = _setup(str(tmp_path), nb, grid, kind)
    N0, N1 = grid
    HR = eng.fourier(-1, o.get("H0k"))
    want = np.zeros_like(HR)
    for n, (i, j, _) in enumerate(m["R"]):
        want[(i % N0) * N1 + (j % N1)] += m["H"][n] / HARTREE_EV
    err = np.abs(HR - want).max() / np.abs(want).max()
    # the mirror image: what a flipped sign would give
    mirror = np.zeros_like(HR)
    for n, (i, j, _) in enumerate(m["R"]):
        mirror[(-i % N0) * N1 + (-j % N1)] += m["H"][n] / HARTREE_EV
    flip = np.abs(HR - mirror).max() / np.abs(want).max()
    _say(nb, grid, len(m["R"]), "R vectors: max|H(R) - tb H(R)| / max =", err, "(mirror image:", flip, ")")
    assert err <= 1e-13
    assert flip > 1e-3


def test_fourier_refuses(tmp_path):
    _, o, eng = _setup(str(tmp_path), 2, [6, 5], "square")
    x = np.zeros((30, 2, 2), dtype=np.complex128)
    with pytest.raises(edus.EdusError, match="sign"):
        eng.fourier(0, x)
    _, _, herm = _setup(str(tmp_path), 2, [6, 5], "square", layout=edus.LAYOUT_HERMITIAN)
    with pytest.raises(edus.EdusError, match="FULL"):
        herm.fourier(-1, x)
