"""output.h5 as the reference's HDF5 build writes it (Simulation.cpp:215-229,
385-420; GlobalFunctions.cpp:19-66; read by the reference's
Postproces/custom_functions/read.py): the /json group, the sparse-print
nodes under /time_au and the density matrices asked for by "toprint",
checked against the input, the sparse print rule (Simulation.cpp:240-257)
and the oracle's state at the printed steps.  Read back with the HDF5 C
library."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import hbn_params

import edus

oracle_py = pytest.importorskip("oracle_py")
h5read = pytest.importorskip("h5read")

pytestmark = pytest.mark.gpu

AU_TIME_FS = 2.4188843265857e-2  # fs per atomic unit of time
HARTREE_EV = 27.211386245988


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_hbn_output_h5(tmp_path):
    try:
        h5read.lib()
    except OSError:
        pytest.skip("no HDF5 library on this machine")
    nsteps = 60
    p = hbn_params(write_h5=True, toprint={"DMk_wannier": "true", "DMk_bloch": "true"},
                   printresolution=4, printresolution_pulse=2)
    sim = edus.Simulation(p)
    sim.write_outputs(str(tmp_path), nsteps)

    # the sparse print rule, restated: resolution 2 inside the pulse, 4 outside
    dt = 0.01 / AU_TIME_FS
    las = p["lasers"][0]
    duration = las["cycles"] * 2 * math.pi / (las["frequency"] / HARTREE_EV)
    steps = []
    for i in range(nsteps):
        t = i * dt
        res = 2 if (t > 1e-7 and t < duration + 1e-7) else 4
        if round(t / dt) % res == 0:
            steps.append(i)
    _say("sparse print steps", steps)
    assert 0 in steps and len(steps) > 5

    o = oracle_py.Oracle(p)
    U = sim.get("U")  # the library's eigenvectors (column phases are the solver's choice)
    ref = {}
    for i in range(nsteps):
        if i in steps:
            ref[i] = o.rho.copy()
        o.step(1)

    nk, nb = o.Nk, o.nb
    with h5read.File(os.path.join(str(tmp_path), "output.h5")) as f:
        js = "/json/"
        assert bytes(f.read(js + "tb_file")).decode() == p["tb_file"]
        assert f.read(js + "coulomb").tolist() == [0]
        assert f.read(js + "grid").tolist() == [50, 50, 1]
        assert f.read(js + "filledbands").tolist() == [1]
        assert f.read(js + "dt").tolist() == [0.01]
        assert f.read(js + "printresolution").tolist() == [4]
        assert f.read(js + "initialtime").tolist() == [0.0] and f.read(js + "finaltime").tolist() == [1.0]
        assert f.read(js + "num_bands").tolist() == [2] and f.read(js + "num_kpoints").tolist() == [2500]
        assert bytes(f.read(js + "solver")).decode() == "RK"
        assert f.read(js + "comm_size").tolist() == [1]
        assert f.read(js + "laser0/intensity").tolist() == [1e10]
        assert f.read(js + "laser0/frequency").tolist() == [10.0]
        assert f.read(js + "laser0/cycles").tolist() == [1]
        assert f.read(js + "laser0/polarization").tolist() == [1]  # the reference writes the cycles there
        kp = f.read(js + "kpoints").reshape(nk, 3)
        i0, i1 = np.divmod(np.arange(nk), 50)
        assert np.array_equal(kp, np.stack([i0 / 50, i1 / 50, np.zeros(nk)], axis=1))
        A = f.read(js + "A").reshape(3, 3)
        B = f.read(js + "B").reshape(3, 3)
        assert np.abs(A - o.lattice()).max() < 1e-12
        assert np.abs(B - 2 * np.pi * np.linalg.inv(A).T).max() < 1e-12
        for g in ("/fullH", "/SelfEnergy"):
            assert f.exists(g) and not f.exists(g + "/0")
        assert not f.exists(f"/time_au/{len(steps)}")
        for node, i in enumerate(steps):
            t = f.read(f"/time_au/{node}")
            assert t.shape == (1,) and abs(t[0] - i * dt) < 1e-9 * max(1.0, i * dt), (node, i, t)
            dm = f.read(f"/DensityMatrix_k/{node}/0/local").view(np.complex128).reshape(nk, nb, nb)
            err = np.abs(dm - ref[i]).max()
            dmb = f.read(f"/DensityMatrix_k_bloch/{node}/0/local").view(np.complex128).reshape(nk, nb, nb)
            bl = U.conj().transpose(0, 2, 1) @ ref[i] @ U
            errb = np.abs(dmb - bl).max()
            if node % 5 == 0:
                _say("  node", node, "step", i, "max|d rho|", err, "bloch", errb)
            assert err <= 1e-10 and errb <= 1e-10


def test_screened_potential_h5_nodes(tmp_path):
    """The screened-potential application's nodes under "-1"
    (Screened_potential.cpp:60-70,104-109,137-142,192-196): Xk is the response
    in k, Uk the eigenvectors, Vk, Epsilonk and Wk the R-space fields taken
    to k with sign +1 and no scaling (numpy's N * ifft2); added to an
    existing output.h5."""
    try:
        h5read.lib()
    except OSError:
        pytest.skip("no HDF5 library on this machine")
    p = hbn_params(write_h5=True, grid=[12, 10, 1])
    sim = edus.Simulation(p)
    sim.write_outputs(str(tmp_path), 3)  # creates output.h5 with the propagation's nodes
    sim.write_screened(str(tmp_path))
    f_ = sim.screened_potential()
    o = oracle_py.Oracle(p)
    nk, nb = o.Nk, o.nb
    Xk_ref = oracle_py.lindhard(sim.get("U"), sim.get("eps"), 12, 10, 1)

    def to_k(a):
        g = a.reshape(12, 10, nb, nb)
        return (np.fft.ifft2(g, axes=(0, 1)) * (12 * 10)).reshape(nk, nb, nb)

    with h5read.File(os.path.join(str(tmp_path), "output.h5")) as f:
        assert f.exists("/json/grid") and f.exists("/time_au/0")  # the earlier content is kept

        def node(name):
            return f.read(f"/-1/{name}/0/local").view(np.complex128).reshape(nk, nb, nb)

        assert np.abs(node("Xk") - Xk_ref).max() <= 1e-11 * np.abs(Xk_ref).max()
        assert np.array_equal(node("Uk"), sim.get("U"))
        for name, key in (("Vk", "V"), ("Epsilonk", "Eps"), ("Wk", "W")):
            ref = to_k(f_[key])
            err = np.abs(node(name) - ref).max()
            _say(name, "max abs err", err, "max", np.abs(ref).max())
            assert err <= 1e-12 * max(1.0, np.abs(ref).max()), name
