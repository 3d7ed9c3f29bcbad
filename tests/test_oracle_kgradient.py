"""CPU: the oracle's k-space finite-difference gradient (gradient_space "k",
kGradient.cpp:25-210, kGradient_definitions.hpp:19-31) pinned before it judges
the GPU.

- B1 condition of the weights: sum_s w_s sum_{b in s} b_a b_b = delta_ab on the
  grid's active directions (what the shell search solves for), and the stencil
  is symmetric (b and -b with equal weight);
- on a square lattice the stencil is the textbook central difference: the four
  nearest mesh vectors with weight 1/(2|b|^2);
- on the hexagonal lattice one shell of six neighbours suffices (w = 1/(3|b|^2));
- a rectangular mesh (unequal spacings) needs a second shell;
- applied to a plane wave F(k) = exp(i k.R0) M it converges to the exact
  derivative i (u.R0) F with second order in the mesh spacing, which pins the
  sign, the k+b neighbour mapping and the scale of b.
"""
import os

import numpy as np
import pytest

import oracle_py
from conftest import hbn_params

from edus import models


def _square(tmp_path, grid):
    return models.workload("ipa2_square30", str(tmp_path / f"sq_{grid[0]}x{grid[1]}"), grid=[grid[0], grid[1], 1])


def _b1(st):
    b = st[:, 3:6]
    w = st[:, 6]
    return np.einsum("i,ia,ib->ab", w, b, b)


def test_hexagonal_one_shell_and_b1():
    o = oracle_py.Oracle(hbn_params(grid=[12, 12, 1], gradient_space="k"))
    st = o.kstencil()
    assert st.shape[0] == 6 and np.all(st[:, 0] == 0)
    bn = np.linalg.norm(st[:, 3:6], axis=1)
    assert np.allclose(bn, bn[0], rtol=1e-12)
    assert np.allclose(st[:, 6], 1.0 / (3.0 * bn[0] ** 2), rtol=1e-12)
    m = _b1(st)
    assert np.allclose(m[:2, :2], np.eye(2), atol=1e-12) and np.allclose(m[2], 0.0)
    # symmetric: every (m0, m1) has its mirror
    offs = {(int(a), int(b)) for a, b in st[:, 1:3]}
    assert all((-a, -b) in offs for a, b in offs)


def test_square_is_central_difference(tmp_path):
    o = oracle_py.Oracle(dict(_square(tmp_path, [8, 8]), gradient_space="k"))
    st = o.kstencil()
    assert st.shape[0] == 4
    offs = sorted((int(a), int(b)) for a, b in st[:, 1:3])
    assert offs == [(-1, 0), (0, -1), (0, 1), (1, 0)]
    bn = np.linalg.norm(st[:, 3:6], axis=1)
    assert np.allclose(st[:, 6], 1.0 / (2.0 * bn ** 2), rtol=1e-12)
    # (u.grad) F along b1 = (F(k+b1) - F(k-b1)) / (2|b1|) for u = b1/|b1|
    rng = np.random.default_rng(0)
    F = rng.standard_normal((o.Nk, 2, 2)) + 1j * rng.standard_normal((o.Nk, 2, 2))
    b1 = st[[i for i in range(4) if (st[i, 1], st[i, 2]) == (1, 0)][0], 3:6]
    u = b1 / np.linalg.norm(b1)
    got = o.kgradient(u, F).reshape(8, 8, 2, 2)
    Fg = F.reshape(8, 8, 2, 2)
    ref = (np.roll(Fg, -1, axis=0) - np.roll(Fg, 1, axis=0)) / (2 * np.linalg.norm(b1))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_rectangular_mesh_needs_two_shells(tmp_path):
    o = oracle_py.Oracle(dict(_square(tmp_path, [8, 12]), gradient_space="k"))
    st = o.kstencil()
    assert set(st[:, 0].astype(int)) == {0, 1}
    m = _b1(st)
    assert np.allclose(m[:2, :2], np.eye(2), atol=1e-12)


@pytest.mark.parametrize("lattice", ["hex", "square"])
def test_plane_wave_second_order(tmp_path, lattice):
    """F(k) = exp(i k.R0) M; exact (u.grad) F = i (u.R0) F."""
    errs = []
    for n in (12, 24, 48):
        if lattice == "hex":
            p = hbn_params(grid=[n, n, 1], gradient_space="k")
        else:
            p = dict(_square(tmp_path, [n, n]), gradient_space="k")
        o = oracle_py.Oracle(p)
        A = o.lattice()  # columns: lattice vectors (bohr)
        R0 = A @ np.array([1.0, -2.0, 0.0])
        B = 2 * np.pi * np.linalg.inv(A).T
        i0, i1 = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        kc = np.stack([i0.ravel() / n, i1.ravel() / n, np.zeros(n * n)], axis=1)
        kcart = kc @ B.T
        phase = np.exp(1j * kcart @ R0)
        Mx = np.array([[1.0, 2.0 - 1j], [0.5j, -1.0]])
        F = phase[:, None, None] * Mx[None]
        u = np.array([0.6, 0.8, 0.0])
        got = o.kgradient(u, F)
        ref = 1j * (u @ R0) * F
        errs.append(np.abs(got - ref).max() / np.abs(ref).max())
    rates = [errs[i] / errs[i + 1] for i in range(2)]
    assert errs[-1] < 2e-2
    assert all(3.5 < r < 4.5 for r in rates), (errs, rates)


@pytest.mark.parametrize("case", ["hex12", "square8", "rect8x12", "hex_odd9x7"])
def test_library_stencil_matches_oracle(tmp_path, case):
    """The library's host-side stencil builder (edus_kgrad_stencil, no GPU)
    against the oracle's: same points in the same order, same weights."""
    import edus

    if case.startswith("hex"):
        grid = [12, 12, 1] if case == "hex12" else [9, 7, 1]
        p = hbn_params(grid=grid, gradient_space="k")
    else:
        grid = [8, 8] if case == "square8" else [8, 12]
        p = dict(_square(tmp_path, grid), gradient_space="k")
    o = oracle_py.Oracle(p)
    ref = o.kstencil()
    got = edus.kgrad_stencil(o.lattice(), o.N0, o.N1)
    assert got.shape == ref.shape
    assert np.array_equal(got[:, :3], ref[:, :3])
    assert np.allclose(got[:, 3:6], ref[:, 3:6], rtol=0, atol=1e-14)
    assert np.allclose(got[:, 6], ref[:, 6], rtol=1e-12)
