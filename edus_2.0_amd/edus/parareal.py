"""Learned parareal (edus_parareal_*): Python side of the C ABI, plus the field
tables and the training loop around it.

The numerics run in the library: F (fine RK4 over a slice, all slices as one
stacked batch), G (coarse RK4 plus the per-k network), the parareal sweeps,
the training-data collection and the network's training kernels.  This module
builds field tables from pulse parameters and the network's normalisation
statistics from the collected data.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .engine import EdusError, Mlp, _check, _dp, _fp, lib

HARTREE_EV = 2 * 2.1798723611035e-18 / 1.602176634e-19
AU_TIME_FS = 2.4188843265857e-2


class Params(C.Structure):
    _fields_ = [("nslices", C.c_int), ("fine_steps", C.c_int), ("coarse_steps", C.c_int), ("dt", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("graph_steps", C.c_int)]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("fallback", C.c_int),
                ("last_update", C.c_double), ("seconds", C.c_double), ("fine_seconds", C.c_double),
                ("coarse_seconds", C.c_double), ("updates", C.c_double * 16)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "updates"}
        d["updates"] = list(self.updates)[:min(self.iterations, 16)]
        return d


def _bind():
    L = lib()
    if getattr(L, "_parareal_bound", False):
        return L
    vp, dp, fp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float)
    L.edus_parareal_create.restype = vp
    L.edus_parareal_create.argtypes = [vp, C.POINTER(Params)]
    L.edus_parareal_destroy.argtypes = [vp]
    L.edus_parareal_set_fields.argtypes = [vp, dp, dp, fp, C.c_int]
    L.edus_parareal_set_network.argtypes = [vp, vp, fp, fp, fp, fp]
    L.edus_parareal_run.argtypes = [vp, dp, C.POINTER(Result)]
    L.edus_parareal_get_state.argtypes = [vp, C.c_int, dp]
    L.edus_parareal_collect.argtypes = [vp, C.POINTER(Params), C.c_int, dp, dp, fp, C.c_int, C.c_int,
                                        C.c_ulonglong, fp, fp]
    L._parareal_bound = True
    return L


# ---------------------------------------------------------------- fields
class Pulse:
    """sin^2 pulse: E(t) = e0 sin^2(pi (t - t0) / T) sin(w (t - t0) + phase) pol
    for 0 < t - t0 <= T, zero outside (atomic units)."""

    def __init__(self, e0: float, omega_ev: float, duration_fs: float, pol=(1.0, 0.0, 0.0), phase: float = 0.0,
                 t0: float = 0.0):
        self.e0, self.w = float(e0), omega_ev / HARTREE_EV
        self.T = duration_fs / AU_TIME_FS
        self.pol = np.asarray(pol, dtype=float)
        self.phase, self.t0 = float(phase), float(t0)

    def __call__(self, t):
        t = np.asarray(t, dtype=float)
        s = t - self.t0
        env = np.where((s > 0) & (s <= self.T), np.sin(math.pi * s / self.T) ** 2, 0.0)
        return (self.e0 * env * np.sin(self.w * s + self.phase))[..., None] * self.pol


def stage_table(field, t0: float, dt: float, nsteps: int) -> np.ndarray:
    """E at t, t + dt/2, t + dt of nsteps RK4 steps from t0: [nsteps][3][3]."""
    t = t0 + dt * np.arange(nsteps)[:, None] + dt * np.array([0.0, 0.5, 1.0])[None, :]
    return np.asarray(field(t), dtype=float).reshape(nsteps, 3, 3)


def slice_tables(field, nslices: int, fine_steps: int, coarse_steps: int, dt: float, t0: float = 0.0,
                 samples: int = 8):
    """Fine and coarse stage tables of every slice and the network's field
    samples (x, y at `samples` points spread over each slice)."""
    T = fine_steps * dt
    dtc = T / coarse_steps
    Ef = np.stack([stage_table(field, t0 + n * T, dt, fine_steps) for n in range(nslices)])
    Ec = np.stack([stage_table(field, t0 + n * T, dtc, coarse_steps) for n in range(nslices)])
    ts = t0 + T * (np.arange(nslices)[:, None] + (np.arange(samples)[None, :] + 0.5) / samples)
    Es = np.asarray(field(ts), dtype=float)[..., :2].astype(np.float32)
    return Ef, Ec, Es


# ---------------------------------------------------------------- driver
class Parareal:
    def __init__(self, engine, nslices: int, fine_steps: int, coarse_steps: int, dt: float, tol: float = 1e-7,
                 max_iter: int = 6, graph_steps: int = 0):
        L = _bind()
        self.engine = engine
        self.p = Params(nslices, fine_steps, coarse_steps, dt, tol, max_iter, graph_steps)
        self.h = L.edus_parareal_create(engine.h, C.byref(self.p))
        if not self.h:
            raise EdusError(-1, L.edus_last_error().decode())
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            lib().edus_parareal_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_fields(self, Ef, Ec, Es):
        ef = np.ascontiguousarray(Ef, dtype=np.float64)
        ec = np.ascontiguousarray(Ec, dtype=np.float64)
        es = np.ascontiguousarray(Es, dtype=np.float32)
        self.field_samples = es.shape[1]
        _check(lib().edus_parareal_set_fields(self.h, _dp(ef), _dp(ec), _fp(es), int(es.shape[1])))

    def set_network(self, net: Mlp | None, stats: dict | None = None):
        if net is None:
            _check(lib().edus_parareal_set_network(self.h, None, None, None, None, None))
            return
        arrs = [np.ascontiguousarray(stats[k], dtype=np.float32) for k in ("mu", "sig", "tmu", "tsig")]
        self._keep = [net] + arrs
        _check(lib().edus_parareal_set_network(self.h, net.h, *[_fp(a) for a in arrs]))

    def run(self, rho0=None) -> dict:
        r = Result()
        if rho0 is None:
            _check(lib().edus_parareal_run(self.h, None, C.byref(r)))
        else:
            a = np.ascontiguousarray(rho0, dtype=np.complex128)
            _check(lib().edus_parareal_run(self.h, _dp(a.view(np.float64)), C.byref(r)))
        return r.as_dict()

    def state(self, n: int) -> np.ndarray:
        e = self.engine
        out = np.empty((e.Nk, e.nb, e.nb), dtype=np.complex128)
        _check(lib().edus_parareal_get_state(self.h, int(n), _dp(out.view(np.float64))))
        return out


def collect(engine, params: Params, Ef, Ec, Es, k_samples: int, seed: int = 0):
    """Training rows from runs r with fields Ef[r], Ec[r], Es[r] (as
    slice_tables returns them): features X [rows][din], targets T [rows][nb^2]."""
    L = _bind()
    ef = np.ascontiguousarray(Ef, dtype=np.float64)
    ec = np.ascontiguousarray(Ec, dtype=np.float64)
    es = np.ascontiguousarray(Es, dtype=np.float32)
    runs, ne = ef.shape[0], es.shape[2]
    nb = engine.nb
    k = min(int(k_samples), engine.Nk)
    din = 2 * nb * nb + 2 * ne + 4
    rows = runs * params.nslices * k
    X = np.empty((rows, din), dtype=np.float32)
    T = np.empty((rows, nb * nb), dtype=np.float32)
    _check(L.edus_parareal_collect(engine.h, C.byref(params), runs, _dp(ef), _dp(ec), _fp(es), ne, k, seed,
                                   _fp(X), _fp(T)))
    return X, T


def normalisation(X, T) -> dict:
    """Feature shift/scale and target scale from training data (a feature or
    target that does not vary keeps scale 1)."""
    mu = X.mean(0)
    sig = X.std(0)
    sig = np.where(sig > 1e-12, sig, 1.0)
    tsig = T.std(0)
    tsig = np.where(tsig > 1e-30, tsig, 1.0)
    return {"mu": mu, "sig": sig, "tmu": np.zeros(T.shape[1], dtype=np.float32), "tsig": tsig}


def train_network(X, T, stats: dict, hidden: int = 256, epochs: int = 40, batch: int = 1024, lr: float = 1e-3,
                  seed: int = 0):
    """The per-k network trained on the GPU on normalised data."""
    Xn = (X - stats["mu"]) / stats["sig"]
    Tn = (T - stats["tmu"]) / stats["tsig"]
    net = Mlp(X.shape[1], hidden, T.shape[1], seed=seed)
    hist = net.train(Xn, Tn, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return net, hist


def random_pulses(n: int, seed: int, e0_range=(1e-3, 1e-2), omega_ev=(0.8, 2.0), duration_fs=(10.0, 40.0)):
    """n pulses with log-uniform peak field, uniform photon energy, CEP and
    in-plane polarisation angle."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        e0 = float(np.exp(rng.uniform(np.log(e0_range[0]), np.log(e0_range[1]))))
        ang = rng.uniform(0, np.pi)
        out.append(Pulse(e0, rng.uniform(*omega_ev), rng.uniform(*duration_fs),
                         (math.cos(ang), math.sin(ang), 0.0), rng.uniform(0, 2 * np.pi)))
    return out
