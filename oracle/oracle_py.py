"""ctypes loader for the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by the smoke check and by bench.py's CPU-baseline leg.
The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Laser(C.Structure):
    _fields_ = [
        ("intensity", C.c_double), ("intensity_units", C.c_char * 32),
        ("frequency", C.c_double), ("frequency_units", C.c_char * 32),
        ("wavelength", C.c_double), ("wavelength_units", C.c_char * 32),
        ("polarization", C.c_double * 3),
        ("cycles", C.c_double),
        ("t0", C.c_double), ("t0_units", C.c_char * 32),
        ("phase", C.c_double),
    ]


class _Params(C.Structure):
    _fields_ = [
        ("tb_file", C.c_char * 1024),
        ("grid", C.c_int * 3),
        ("filledbands", C.c_int),
        ("dt", C.c_double), ("dt_units", C.c_char * 32),
        ("initialtime", C.c_double), ("initialtime_units", C.c_char * 32),
        ("finaltime", C.c_double), ("finaltime_units", C.c_char * 32),
        ("printresolution", C.c_int), ("printresolution_pulse", C.c_int),
        ("solver", C.c_char * 8), ("order", C.c_int),
        ("coulomb", C.c_int), ("epsilon", C.c_double),
        ("nr0", C.c_int), ("r0", C.c_double * 3), ("r0_units", C.c_char * 32),
        ("coulomb_center_quirk", C.c_int),
        ("nlasers", C.c_int),
        ("lasers", _Laser * 10),
        ("peierls", C.c_int),
        ("gradient_k", C.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [C.POINTER(_Params)]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_destroy.argtypes = [vp]
        L.oracle_dims.argtypes = [vp, ip, ip, ip, ip, dp]
        L.oracle_time.restype = C.c_double
        L.oracle_time.argtypes = [vp]
        L.oracle_get.restype = C.c_long
        L.oracle_get.argtypes = [vp, C.c_char_p, dp]
        L.oracle_set.argtypes = [vp, C.c_char_p, dp]
        L.oracle_eigenvalues.argtypes = [vp, dp]
        L.oracle_rhs.argtypes = [vp, C.c_double, dp, dp]
        L.oracle_rhs_E.argtypes = [vp, dp, dp, dp]
        L.oracle_rhs_EA.argtypes = [vp, dp, dp, dp, dp]
        L.oracle_laser_A.argtypes = [vp, dp]
        L.oracle_laser_A.restype = C.c_int
        L.oracle_kstencil.argtypes = [vp, dp]
        L.oracle_kstencil.restype = C.c_int
        L.oracle_kgradient.argtypes = [vp, dp, dp, dp]
        L.oracle_kgradient.restype = C.c_int
        L.oracle_vector_potential.argtypes = [vp, C.c_double, dp]
        L.oracle_lattice.argtypes = [vp, dp]
        L.oracle_step.argtypes = [vp, C.c_int]
        L.oracle_E.argtypes = [vp, C.c_double, dp]
        L.oracle_observables.argtypes = [vp, dp, dp, dp]
        L.oracle_rhoR.argtypes = [vp, dp]
        L.oracle_lindhard.restype = C.c_int
        L.oracle_lindhard.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.oracle_screened.argtypes = [vp, dp, dp, dp, dp, dp]
        L.oracle_rwann.argtypes = [vp, dp]
        L.oracle_propagate.restype = C.c_int
        L.oracle_propagate.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp]
        L.oracle_eigh.argtypes = [C.c_int, dp, dp, dp]
        L.oracle_fft2d.argtypes = [C.c_int, C.c_int, dp, C.c_int]
        L.oracle_struve0.restype = C.c_double
        L.oracle_struve0.argtypes = [C.c_double]
        L.oracle_rotate.argtypes = [vp, C.c_int, dp, dp]
        L.oracle_host_bandwidth.restype = C.c_double
        L.oracle_host_bandwidth.argtypes = [C.c_long, C.c_int]
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(p: dict) -> _Params:
    s = _Params()
    s.tb_file = p["tb_file"].encode()
    for i in range(3):
        s.grid[i] = int(p["grid"][i])
    s.filledbands = int(p["filledbands"])
    for key in ("dt", "initialtime", "finaltime"):
        setattr(s, key, float(p[key]))
        setattr(s, key + "_units", p[key + "_units"].encode())
    s.printresolution = int(p["printresolution"])
    s.printresolution_pulse = int(p["printresolution_pulse"])
    s.solver = p["solver"].encode()
    s.order = int(p["order"])
    s.coulomb = int(bool(p["coulomb"]))
    s.epsilon = float(p["epsilon"])
    r0 = list(p["r0"])
    s.nr0 = len(r0)
    for i, v in enumerate(r0[:3]):
        s.r0[i] = float(v)
    s.r0_units = p["r0_units"].encode()
    s.coulomb_center_quirk = int(bool(p.get("coulomb_center_quirk", True)))
    s.peierls = int(bool(p.get("peierls", False)))
    s.gradient_k = int(p.get("gradient_space", "R") == "k")
    s.nlasers = len(p["lasers"])
    for i, las in enumerate(p["lasers"]):
        L = s.lasers[i]
        for key in ("intensity", "frequency", "wavelength", "t0"):
            setattr(L, key, float(las[key]))
            setattr(L, key + "_units", las[key + "_units"].encode())
        for d in range(3):
            L.polarization[d] = float(las["polarization"][d])
        L.cycles = float(las["cycles"])
        L.phase = float(las["phase"])
    return s


class Oracle:
    """The reference's Simulation, restated on the CPU in FP64."""

    def __init__(self, params: dict):
        L = lib()
        self._p = make_params(params)
        self.h = L.oracle_create(C.byref(self._p))
        if not self.h:
            raise RuntimeError(L.oracle_last_error().decode())
        nb, n0, n1, ns = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        dt = C.c_double()
        L.oracle_dims(self.h, C.byref(nb), C.byref(n0), C.byref(n1), C.byref(ns), C.byref(dt))
        self.nb, self.N0, self.N1, self.nsteps, self.dt = nb.value, n0.value, n1.value, ns.value, dt.value
        self.Nk = self.N0 * self.N1

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def get(self, name: str) -> np.ndarray:
        L = lib()
        n = L.oracle_get(self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=np.complex128)
        L.oracle_get(self.h, name.encode(), _ptr(out.view(np.float64)))
        if name == "hartree":
            return out.reshape(self.nb, self.nb)
        return out.reshape(-1, self.nb, self.nb)

    def set(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        if lib().oracle_set(self.h, name.encode(), _ptr(a.view(np.float64))) != 0:
            raise KeyError(name)

    @property
    def rho(self) -> np.ndarray:
        return self.get("rho")

    @property
    def time(self) -> float:
        return lib().oracle_time(self.h)

    def eigenvalues(self) -> np.ndarray:
        out = np.empty(self.Nk * self.nb)
        lib().oracle_eigenvalues(self.h, _ptr(out))
        return out.reshape(self.Nk, self.nb)

    def rhs(self, t: float, rho: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(rho, dtype=np.complex128)
        out = np.empty_like(a)
        lib().oracle_rhs(self.h, float(t), _ptr(a.view(np.float64)), _ptr(out.view(np.float64)))
        return out

    def rotate(self, to_bloch: bool, x: np.ndarray) -> np.ndarray:
        """Operator::go_to_bloch (U^dagger x U) or go_to_wannier (U x U^dagger) per k."""
        a = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(a)
        lib().oracle_rotate(self.h, int(bool(to_bloch)), _ptr(a.view(np.float64)), _ptr(out.view(np.float64)))
        return out

    def rhs_E(self, E, rho: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(rho, dtype=np.complex128)
        e = np.ascontiguousarray(E, dtype=np.float64)
        out = np.empty_like(a)
        lib().oracle_rhs_E(self.h, _ptr(e), _ptr(a.view(np.float64)), _ptr(out.view(np.float64)))
        return out

    def rhs_EA(self, E, Av, rho: np.ndarray) -> np.ndarray:
        """Peierls-gauge right-hand side with field E and vector potential Av."""
        a = np.ascontiguousarray(rho, dtype=np.complex128)
        e = np.ascontiguousarray(E, dtype=np.float64)
        av = np.ascontiguousarray(Av, dtype=np.float64)
        out = np.empty_like(a)
        lib().oracle_rhs_EA(self.h, _ptr(e), _ptr(av), _ptr(a.view(np.float64)), _ptr(out.view(np.float64)))
        return out

    def kstencil(self) -> np.ndarray:
        """Finite-difference gradient stencil: rows (shell, m0, m1, bx, by, bz, w)."""
        n = lib().oracle_kstencil(self.h, None)
        out = np.empty((n, 7))
        if n:
            lib().oracle_kstencil(self.h, _ptr(out))
        return out

    def kgradient(self, u, F: np.ndarray) -> np.ndarray:
        """(u . grad_k) F by the finite-difference stencil (gradient_space "k")."""
        a = np.ascontiguousarray(F, dtype=np.complex128)
        uu = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty_like(a)
        if lib().oracle_kgradient(self.h, _ptr(uu), _ptr(a.view(np.float64)), _ptr(out.view(np.float64))) != 0:
            raise RuntimeError("oracle: gradient_space is not k")
        return out

    def vector_potential(self, t: float) -> np.ndarray:
        out = np.empty(3)
        lib().oracle_vector_potential(self.h, float(t), _ptr(out))
        return out

    def lattice(self) -> np.ndarray:
        out = np.empty(9)
        lib().oracle_lattice(self.h, _ptr(out))
        return out.reshape(3, 3)

    def step(self, n: int = 1):
        lib().oracle_step(self.h, int(n))

    def E(self, t: float) -> np.ndarray:
        out = np.empty(3)
        lib().oracle_E(self.h, float(t), _ptr(out))
        return out

    def observables(self):
        pb, pw, v = np.empty(self.nb), np.empty(self.nb), np.empty(6)
        lib().oracle_observables(self.h, _ptr(pb), _ptr(pw), _ptr(v))
        return pb, pw, v

    def rhoR(self) -> np.ndarray:
        out = np.empty((self.Nk, self.nb, self.nb), dtype=np.complex128)
        lib().oracle_rhoR(self.h, _ptr(out.view(np.float64)))
        return out

    def screened_potential(self) -> dict:
        """ScreenedPotential::Calculate: R-space fields X, V, Eps, InvEps, W [R][n][m]."""
        out = {k: np.empty((self.Nk, self.nb, self.nb), dtype=np.complex128) for k in ("X", "V", "Eps", "InvEps", "W")}
        lib().oracle_screened(self.h, *(_ptr(out[k].view(np.float64)) for k in ("X", "V", "Eps", "InvEps", "W")))
        return out

    def wannier_centres(self) -> np.ndarray:
        """Diagonal of r(R = 0) per orbital (bohr), [nb][3]."""
        out = np.empty((self.nb, 3))
        lib().oracle_rwann(self.h, _ptr(out))
        return out

    def propagate(self, nsteps: int | None = None) -> dict:
        n = self.nsteps if nsteps is None else int(nsteps)
        times, laser = np.empty(n), np.empty((n, 3))
        pb, pw, vel = np.empty((n, self.nb)), np.empty((n, self.nb)), np.empty((n, 6))
        k = lib().oracle_propagate(self.h, n, _ptr(times), _ptr(laser), _ptr(pb), _ptr(pw), _ptr(vel))
        la = np.empty((max(k, 1), 3))
        lib().oracle_laser_A(self.h, _ptr(la))
        return {"time": times[:k], "laser": laser[:k], "pop_bloch": pb[:k], "pop_wannier": pw[:k],
                "velocity": vel[:k], "laser_A": la[:k]}


def eigh(A: np.ndarray):
    n = A.shape[0]
    a = np.ascontiguousarray(A, dtype=np.complex128)
    w = np.empty(n)
    V = np.empty((n, n), dtype=np.complex128)
    lib().oracle_eigh(n, _ptr(a.view(np.float64)), _ptr(w), _ptr(V.view(np.float64)))
    return w, V


def fft2d(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.array(x, dtype=np.complex128, order="C", copy=True)
    lib().oracle_fft2d(a.shape[0], a.shape[1], _ptr(a.view(np.float64)), int(sign))
    return a


def struve0(x: float) -> float:
    return lib().oracle_struve0(float(x))


def lindhard(U: np.ndarray, eps: np.ndarray, N0: int, N1: int, filled: int) -> np.ndarray:
    """Lindhard response X(q) [q][a][b] (Screened_potential.cpp:17-58) for
    eigenvectors U [k][a][n] and energies eps [k][n] on an N0 x N1 grid."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    nk, nb, _ = U.shape
    assert nk == N0 * N1 and e.shape == (nk, nb)
    X = np.empty_like(U)
    if lib().oracle_lindhard(nb, N0, N1, int(filled), _ptr(U.view(np.float64)), _ptr(e), _ptr(X.view(np.float64))) != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return X


def host_bandwidth(nbytes: int = 1 << 30, reps: int = 3) -> float:
    """Host memory bandwidth (bytes read + written per second) of an OpenMP
    parallel copy between two arrays of nbytes each."""
    return float(lib().oracle_host_bandwidth(int(nbytes), int(reps)))
