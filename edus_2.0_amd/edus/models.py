"""Seeded synthetic tight-binding models written in the reference's
Wannier90 "_tb.dat" format (parsed by src/Wannier/ReadWannier.cpp:110-153), and
reference-key simulation inputs for the benchmark workloads.

Workloads (dt = 0.05 a.u. throughout):
  ipa2_square30     2-band square lattice, 30x30 k, RK4, x-polarised 1.5 eV, 0.01 a.u., 10 fs
  ipa14_hex120      14-band hexagonal model, 120x120 k, RK4, 1.0 eV, 0.02 a.u., 20 fs
  hsex8_hex90       8-band hexagonal HSEX model, 90x90 k, AB4, circular 1.8 eV, 0.005 a.u., 15 fs
  hsex8_hex96       8-band hexagonal HSEX model, 96x96 k, RK4 (parareal workload)
  ipa6_hex96        6-band hexagonal model, 96x96 k, RK4 (a batch member class)
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import config

BOHR_ANGSTROM = 0.52917721090380
HARTREE_EV = 2 * 2.1798723611035e-18 / 1.602176634e-19
AU_TIME_FS = 2.4188843265857e-17 / 1e-15


def lattice_vectors(kind: str, a: float) -> np.ndarray:
    """Rows are lattice vectors in angstrom (vacuum of 20 A along z)."""
    if kind == "square":
        return np.array([[a, 0, 0], [0, a, 0], [0, 0, 20.0]])
    if kind == "hex":
        return np.array([[a * math.sqrt(3) / 2, a / 2, 0], [a * math.sqrt(3) / 2, -a / 2, 0], [0, 0, 20.0]])
    raise ValueError(kind)


def _rvectors(uc: np.ndarray, rmax: float):
    a = np.linalg.norm(uc[0])
    n = int(math.ceil(rmax)) + 2
    out = []
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            r = np.linalg.norm(i * uc[0] + j * uc[1]) / a
            if r <= rmax + 1e-9:
                out.append((i, j, 0, r))
    out.sort(key=lambda t: (t[3], t[0], t[1]))
    return out


def random_model(nb: int, kind: str, rmax: float, amp_ev: float, onsite_ev, xi0_bohr: float,
                 xi_decay_bohr: float, seed: int, a: float = 2.5, centres=None, soc: float = 0.0) -> dict:
    """Random Hermitian H(R) ~ amp e^{-|R|} for |R| <= rmax (lattice units),
    H(-R) = H(R)^dagger; onsite energies on the R=0 diagonal; position matrix
    with random off-diagonal entries (xi0 at R=0, xi_decay e^{-|R|} elsewhere)
    and Wannier centres on its R=0 diagonal.  soc > 0 adds a spin-orbit-like
    2x2 block coupling between consecutive orbital pairs."""
    rng = np.random.default_rng(seed)
    uc = lattice_vectors(kind, a)
    Rs = _rvectors(uc, rmax)
    index = {(i, j, k): n for n, (i, j, k, _) in enumerate(Rs)}
    nR = len(Rs)
    H = np.zeros((nR, nb, nb), dtype=np.complex128)
    X = np.zeros((3, nR, nb, nb), dtype=np.complex128)
    xi0 = xi0_bohr * BOHR_ANGSTROM
    xid = xi_decay_bohr * BOHR_ANGSTROM

    def rand(shape, scale):
        return scale * (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))

    for n, (i, j, k, r) in enumerate(Rs):
        m = index[(-i, -j, -k)]
        if m < n:
            continue
        if n == m:  # R = 0
            h = rand((nb, nb), amp_ev)
            h = 0.5 * (h + h.conj().T)
            h[np.diag_indices(nb)] = np.asarray(onsite_ev, dtype=float)
            H[n] = h
            for d in range(2):
                x = rand((nb, nb), xi0)
                x = 0.5 * (x + x.conj().T)
                x[np.diag_indices(nb)] = 0.0
                X[d, n] = x
        else:
            h = rand((nb, nb), amp_ev * math.exp(-r))
            if soc > 0 and nb % 2 == 0:
                for b in range(0, nb, 2):
                    h[b, b + 1] += soc * amp_ev * math.exp(-r) * 1j
                    h[b + 1, b] += soc * amp_ev * math.exp(-r) * 1j
            H[n] = h
            H[m] = h.conj().T
            for d in range(2):
                x = rand((nb, nb), xid * math.exp(-r))
                X[d, n] = x
                X[d, m] = x.conj().T
    if centres is not None:
        c = np.asarray(centres, dtype=float)
        n0 = index[(0, 0, 0)]
        for d in range(3):
            X[d, n0][np.diag_indices(nb)] = c[:, d]
    return {"uc": uc, "R": [(i, j, k) for (i, j, k, _) in Rs], "H": H, "r": X, "nb": nb}


def write_tb(path: str, model: dict) -> None:
    """Write a _tb.dat file (H in eV, r in angstrom, all ndegen = 1)."""
    nb, Rs, H, X = model["nb"], model["R"], model["H"], model["r"]
    with open(path, "w") as f:
        f.write("synthetic seeded model\n")
        for row in model["uc"]:
            f.write(f"{row[0]:24.16f} {row[1]:24.16f} {row[2]:24.16f}\n")
        f.write(f"{nb:12d}\n{len(Rs):12d}\n")
        for i in range(len(Rs)):
            f.write(f"{1:5d}")
            if i % 15 == 14 or i == len(Rs) - 1:
                f.write("\n")
        for n, (i, j, k) in enumerate(Rs):
            f.write(f"\n{i:5d}{j:5d}{k:5d}\n")
            for b in range(nb):
                for a in range(nb):
                    v = H[n, a, b]
                    f.write(f"{a + 1:5d}{b + 1:5d} {v.real:24.16e} {v.imag:24.16e}\n")
        for n, (i, j, k) in enumerate(Rs):
            f.write(f"\n{i:5d}{j:5d}{k:5d}\n")
            for b in range(nb):
                for a in range(nb):
                    f.write(f"{a + 1:5d}{b + 1:5d}")
                    for d in range(3):
                        v = X[d, n, a, b]
                        f.write(f" {v.real:24.16e} {v.imag:24.16e}")
                    f.write("\n")


def laser(e0_au: float, omega_ev: float, duration_fs: float, pol=(1, 0, 0), phase: float = 0.0) -> dict:
    period_fs = 2 * math.pi / (omega_ev / HARTREE_EV) * AU_TIME_FS
    return {"intensity": e0_au ** 2, "intensity_units": "auintensity",
            "frequency": omega_ev, "frequency_units": "electronvolt",
            "polarization": list(pol), "cycles": duration_fs / period_fs,
            "t0": 0.0, "t0_units": "fs", "phase": phase}


def workload(name: str, outdir: str, grid=None, seed=None) -> dict:
    """Write the model of a named workload under outdir and return its
    composed parameter dict (reference keys)."""
    os.makedirs(outdir, exist_ok=True)
    if name == "ipa2_square30":
        m = random_model(2, "square", 2.0, 0.5, [-1.0, 1.0], 0.1, 0.02, 0 if seed is None else seed, a=3.0)
        user = {"grid": grid or [30, 30, 1], "filledbands": 1, "finaltime": 10.0, "finaltime_units": "fs",
                "lasers": [laser(0.01, 1.5, 10.0, (1, 0, 0))]}
    elif name == "ipa14_hex120":
        ons = list(np.linspace(-3.0, -1.0, 7)) + list(np.linspace(1.0, 3.0, 7))
        m = random_model(14, "hex", 3.0, 1.0, ons, 0.05, 0.05, 1 if seed is None else seed, soc=0.2)
        user = {"grid": grid or [120, 120, 1], "filledbands": 7, "finaltime": 20.0, "finaltime_units": "fs",
                "lasers": [laser(0.02, 1.0, 20.0, (1, 0, 0))]}
    elif name == "ipa6_hex96":
        ons = list(np.linspace(-2.0, -1.0, 3)) + list(np.linspace(1.0, 2.0, 3))
        m = random_model(6, "hex", 3.0, 1.0, ons, 0.05, 0.05, 5 if seed is None else seed)
        user = {"grid": grid or [96, 96, 1], "filledbands": 3, "finaltime": 10.0, "finaltime_units": "fs",
                "lasers": [laser(0.01, 1.2, 10.0, (1, 0, 0))]}
    elif name in ("hsex8_hex90", "hsex8_hex96"):
        ons = list(np.linspace(-3.0, -1.5, 4)) + list(np.linspace(1.5, 3.0, 4))
        a = 2.5
        # four orbitals on each of two sites (a honeycomb-like basis)
        site_b = np.array([a / math.sqrt(3), 0.0, 0.0])
        centres = [np.zeros(3)] * 4 + [site_b] * 4
        seed_ = (2 if name.endswith("90") else 4) if seed is None else seed
        m = random_model(8, "hex", 3.0, 1.0, ons, 0.05, 0.05, seed_, a=a, centres=centres)
        if name.endswith("90"):
            user = {"grid": grid or [90, 90, 1], "filledbands": 4, "finaltime": 15.0, "finaltime_units": "fs",
                    "solver": "AB", "order": 4,
                    "lasers": [laser(0.005, 1.8, 15.0, (1, 0, 0), 0.0),
                               laser(0.005, 1.8, 15.0, (0, 1, 0), math.pi / 2)]}
        else:
            user = {"grid": grid or [96, 96, 1], "filledbands": 4, "finaltime": 320 * AU_TIME_FS,
                    "finaltime_units": "fs", "lasers": [laser(0.005, 1.8, 40.0, (1, 0, 0))]}
        # the reference's centre accumulation is not Hermitian for two off-origin sites
        user.update({"coulomb": True, "epsilon": 1.0, "r0": [40.0], "r0_units": "angstrom",
                     "coulomb_center_quirk": False})
    else:
        raise KeyError(name)
    tb = os.path.join(outdir, name)
    write_tb(tb + "_tb.dat", m)
    user.update({"tb_file": tb, "dt": 0.05, "dt_units": "autime"})
    return config.compose(user)
