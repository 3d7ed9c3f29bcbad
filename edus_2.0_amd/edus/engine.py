"""ctypes binding of the C ABI in include/edus_gpu.h (libedus_gpu.so).

The library is built in-tree (`make -C edus_2.0_amd`); nothing here falls back
to another implementation: if the library or a GPU is missing the calls fail.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libedus_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "edus_gpu.h")

PREC_FP64 = 0
PREC_MIXED = 1
LAYOUT_FULL = 0
LAYOUT_HERMITIAN = 1

_LIB = None


class EdusError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"edus error {code}: {msg}")
        self.code = code


class _Desc(C.Structure):
    _fields_ = [
        ("nb", C.c_int), ("N0", C.c_int), ("N1", C.c_int), ("precision", C.c_int), ("layout", C.c_int),
        ("H0k", C.c_void_p), ("xik", C.c_void_p * 3),
        ("A", (C.c_double * 3) * 3),
        ("coulomb", C.c_int), ("W", C.c_void_p), ("hartree", C.c_void_p), ("rho0k", C.c_void_p),
        ("U", C.c_void_p), ("vel", C.c_void_p * 3),
        ("filledbands", C.c_int), ("peierls", C.c_int), ("gradient_space", C.c_int),
    ]


class Tuning(C.Structure):
    """edus_tuning: kernel and tile choices of a context."""
    _fields_ = [("fft_family", C.c_int * 2), ("fft_cfg", C.c_int * 2), ("fft_lines", C.c_int * 2),
                ("comm_kind", C.c_int), ("comm_tile", C.c_int), ("comm_tile_ab", C.c_int), ("grad_fused", C.c_int),
                ("tri_fields", C.c_int), ("hf_fused", C.c_int)]

    def as_dict(self):
        return {"fft_family": list(self.fft_family), "fft_cfg": list(self.fft_cfg),
                "fft_lines": list(self.fft_lines), "comm_kind": int(self.comm_kind),
                "comm_tile": int(self.comm_tile), "comm_tile_ab": int(self.comm_tile_ab),
                "grad_fused": int(self.grad_fused), "tri_fields": int(self.tri_fields),
                "hf_fused": int(self.hf_fused)}


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


class _SimParams(C.Structure):
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
        ("precision", C.c_int),
        ("layout", C.c_int), ("ab_seed", C.c_int), ("peierls", C.c_int), ("gradient_space", C.c_int),
        ("write_h5", C.c_int), ("toprint_dmk_wannier", C.c_int), ("toprint_dmk_bloch", C.c_int),
    ]


EXPORTS = [
    "edus_create", "edus_clone", "edus_destroy", "edus_last_error", "edus_set_state", "edus_get_state", "edus_rhs",
    "edus_rotate", "edus_get_state_bloch", "edus_get_state_bloch_device", "edus_monitor_limits", "edus_fourier",
    "edus_ctx_info",
    "edus_rk4_steps", "edus_ab_steps", "edus_ab_reset", "edus_ab_seed", "edus_rhs_peierls",
    "edus_rk4_steps_peierls", "edus_ab_steps_peierls", "edus_observables_peierls", "edus_observables", "edus_check", "edus_synchronize", "edus_graph",
    "edus_timing_enable", "edus_timing_get", "edus_get_tuning", "edus_set_tuning", "edus_kgrad_stencil",
    "edus_batch_create",
    "edus_batch_destroy", "edus_batch_groups", "edus_batch_graph", "edus_batch_rk4_steps", "edus_batch_store",
    "edus_batch_load", "edus_batch_observables", "edus_batch_check", "edus_batch_synchronize", "edus_mlp_create",
    "edus_mlp_destroy", "edus_mlp_num_params", "edus_mlp_set_params", "edus_mlp_get_params", "edus_mlp_forward",
    "edus_mlp_gradients", "edus_mlp_train", "edus_sim_create", "edus_sim_destroy", "edus_sim_info",
    "edus_sim_engine", "edus_sim_field", "edus_sim_time", "edus_sim_propagate", "edus_sim_steps", "edus_sim_write_outputs", "edus_sim_get",
    "edus_lindhard", "edus_lindhard_device", "edus_lindhard_device_noalloc", "edus_sim_screened", "edus_sim_write_screened",
]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `make -C edus_2.0_amd` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.edus_create.restype = vp
        L.edus_create.argtypes = [C.POINTER(_Desc)]
        L.edus_destroy.argtypes = [vp]
        L.edus_clone.restype = vp
        L.edus_clone.argtypes = [vp]
        L.edus_last_error.restype = C.c_char_p
        for name in ("edus_set_state", "edus_get_state"):
            getattr(L, name).argtypes = [vp, dp]
        L.edus_rhs.argtypes = [vp, dp, dp, dp]
        L.edus_rotate.argtypes = [vp, C.c_int, dp, dp]
        L.edus_fourier.argtypes = [vp, C.c_int, dp, dp]
        ip = C.POINTER(C.c_int)
        L.edus_ctx_info.argtypes = [vp, ip, ip, ip, ip]
        L.edus_get_state_bloch.argtypes = [vp, dp]
        L.edus_get_state_bloch_device.argtypes = [vp, vp]
        L.edus_rk4_steps.argtypes = [vp, C.c_double, C.c_int, dp]
        L.edus_ab_steps.argtypes = [vp, C.c_double, C.c_int, C.c_int, dp]
        L.edus_ab_reset.argtypes = [vp]
        L.edus_ab_seed.argtypes = [vp, C.c_double, C.c_int, dp]
        L.edus_rhs_peierls.argtypes = [vp, dp, dp, dp, dp]
        L.edus_rk4_steps_peierls.argtypes = [vp, C.c_double, C.c_int, dp, dp]
        L.edus_ab_steps_peierls.argtypes = [vp, C.c_double, C.c_int, C.c_int, dp, dp]
        L.edus_observables_peierls.argtypes = [vp, dp, dp, dp, dp]
        L.edus_observables.argtypes = [vp, dp, dp, dp]
        L.edus_check.argtypes = [vp, dp, dp]
        L.edus_monitor_limits.argtypes = [vp, C.c_double, C.c_double]
        L.edus_synchronize.argtypes = [vp]
        L.edus_graph.argtypes = [vp, C.c_int]
        L.edus_timing_enable.argtypes = [vp, C.c_int]
        L.edus_timing_get.argtypes = [vp, C.c_char_p, dp, lp]
        L.edus_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        fp = C.POINTER(C.c_float)
        L.edus_mlp_create.restype = vp
        L.edus_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong]
        L.edus_mlp_destroy.argtypes = [vp]
        L.edus_mlp_num_params.argtypes = [vp, lp]
        L.edus_mlp_set_params.argtypes = [vp, fp]
        L.edus_mlp_get_params.argtypes = [vp, fp]
        L.edus_mlp_forward.argtypes = [vp, C.c_int, fp, fp]
        L.edus_mlp_gradients.argtypes = [vp, C.c_int, fp, fp, fp, fp]
        L.edus_mlp_train.argtypes = [vp, C.c_int, fp, fp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp]
        L.edus_batch_create.restype = vp
        L.edus_batch_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.edus_batch_destroy.argtypes = [vp]
        L.edus_batch_groups.argtypes = [vp, ip, ip]
        L.edus_batch_graph.argtypes = [vp, C.c_int]
        L.edus_batch_rk4_steps.argtypes = [vp, C.c_double, C.c_int, dp]
        L.edus_batch_store.argtypes = [vp, C.c_int]
        L.edus_batch_load.argtypes = [vp, C.c_int]
        L.edus_batch_observables.argtypes = [vp, C.c_int, dp, dp, dp]
        L.edus_batch_check.argtypes = [vp, dp, dp]
        L.edus_batch_synchronize.argtypes = [vp]
        L.edus_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.edus_kgrad_stencil.argtypes = [C.POINTER(C.c_double * 9), C.c_int, C.c_int, dp, C.c_int]
        L.edus_kgrad_stencil.restype = C.c_int
        L.edus_sim_create.restype = vp
        L.edus_sim_create.argtypes = [C.POINTER(_SimParams)]
        L.edus_sim_destroy.argtypes = [vp]
        L.edus_sim_info.argtypes = [vp, ip, ip, ip, ip, dp]
        L.edus_sim_engine.restype = vp
        L.edus_sim_engine.argtypes = [vp]
        L.edus_sim_field.argtypes = [vp, C.c_double, dp]
        L.edus_sim_time.restype = C.c_double
        L.edus_sim_time.argtypes = [vp]
        L.edus_sim_propagate.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp, ip]
        L.edus_sim_steps.argtypes = [vp, C.c_int]
        L.edus_sim_write_outputs.argtypes = [vp, C.c_char_p, C.c_int, ip]
        L.edus_sim_get.restype = C.c_long
        L.edus_sim_get.argtypes = [vp, C.c_char_p, dp]
        L.edus_lindhard.restype = C.c_int
        L.edus_lindhard.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.edus_lindhard_device.restype = C.c_int
        L.edus_lindhard_device.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.edus_lindhard_device_noalloc.restype = C.c_int
        L.edus_lindhard_device_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
        L.edus_sim_screened.restype = C.c_int
        L.edus_sim_screened.argtypes = [vp, dp, dp, dp, dp, dp]
        L.edus_sim_write_screened.restype = C.c_int
        L.edus_sim_write_screened.argtypes = [vp, C.c_char_p]
        _LIB = L
    return _LIB


def _check(rc: int):
    if rc != 0:
        raise EdusError(rc, lib().edus_last_error().decode())


def _dp(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


class Engine:
    """Low-level engine context (edus_create) built from set-up arrays."""

    def __init__(self, nb, N0, N1, H0k, xik=(None, None, None), A=None, filledbands=0, U=None,
                 vel=(None, None, None), coulomb=False, W=None, hartree=None, rho0k=None,
                 precision=PREC_FP64, layout=LAYOUT_FULL, peierls=False, gradient_space="R", _handle=None,
                 _owner=None):
        self.nb, self.N0, self.N1 = nb, N0, N1
        self.Nk = N0 * N1
        self._owner = _owner
        if _handle is not None:
            self.h = _handle
            self._owned = False
            return
        self._keep = []
        d = _Desc()
        d.nb, d.N0, d.N1, d.precision, d.layout = nb, N0, N1, precision, layout

        def ptr(x):
            if x is None:
                return None
            a = _c128(x)
            self._keep.append(a)
            return a.ctypes.data

        d.H0k = ptr(H0k)
        for i in range(3):
            d.xik[i] = ptr(xik[i])
            d.vel[i] = ptr(vel[i])
        A = np.eye(3) if A is None else np.asarray(A, dtype=float)
        for i in range(3):
            for j in range(3):
                d.A[i][j] = A[i, j]
        d.coulomb = int(bool(coulomb))
        d.W, d.hartree, d.rho0k, d.U = ptr(W), ptr(hartree), ptr(rho0k), ptr(U)
        d.filledbands = int(filledbands)
        d.peierls = int(bool(peierls))
        if gradient_space not in ("R", "k"):
            raise ValueError("gradient_space must be 'R' or 'k'")
        d.gradient_space = 1 if gradient_space == "k" else 0
        self.h = lib().edus_create(C.byref(d))
        if not self.h:
            raise EdusError(-1, lib().edus_last_error().decode())
        self._owned = True

    def close(self):
        if getattr(self, "_owned", False) and self.h:
            lib().edus_destroy(self.h)
        self.h = None

    def clone(self) -> "Engine":
        """edus_clone: an independent context with copies of this one's arrays and state."""
        h = lib().edus_clone(self.h)
        if not h:
            raise EdusError(-1, lib().edus_last_error().decode())
        e = Engine(self.nb, self.N0, self.N1, None, _handle=h)
        e._owned = True
        return e

    def __del__(self):
        self.close()

    def _shape(self):
        return (self.Nk, self.nb, self.nb)

    def set_state(self, rho):
        a = _c128(rho).reshape(self._shape())
        _check(lib().edus_set_state(self.h, _dp(a.view(np.float64))))

    def get_state(self) -> np.ndarray:
        out = np.empty(self._shape(), dtype=np.complex128)
        _check(lib().edus_get_state(self.h, _dp(out.view(np.float64))))
        return out

    def rotate(self, to_bloch: bool, x) -> np.ndarray:
        """U^dagger x U (to_bloch) or U x U^dagger per k (edus_rotate)."""
        a = _c128(x).reshape(self._shape())
        out = np.empty_like(a)
        _check(lib().edus_rotate(self.h, int(bool(to_bloch)), _dp(a.view(np.float64)), _dp(out.view(np.float64))))
        return out

    def fourier(self, sign: int, x) -> np.ndarray:
        """k -> R (sign -1, scaled by 1/Nk) or R -> k (sign +1) of an operator
        field [Nk][nb][nb] on the context's grid (edus_fourier)."""
        a = _c128(x).reshape(self._shape())
        out = np.empty_like(a)
        _check(lib().edus_fourier(self.h, int(sign), _dp(a.view(np.float64)), _dp(out.view(np.float64))))
        return out

    def info(self) -> tuple:
        """(nb, N0, N1, layout) of the context (edus_ctx_info)."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().edus_ctx_info(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def get_state_bloch(self) -> np.ndarray:
        out = np.empty(self._shape(), dtype=np.complex128)
        _check(lib().edus_get_state_bloch(self.h, _dp(out.view(np.float64))))
        return out

    def get_state_bloch_device(self, ptr: int) -> None:
        """edus_get_state_bloch_device into the device buffer at ptr (Nk*nb*nb
        complex doubles), queued on the context's stream."""
        _check(lib().edus_get_state_bloch_device(self.h, C.c_void_p(ptr)))

    def rhs(self, E, rho) -> np.ndarray:
        a = _c128(rho).reshape(self._shape())
        e = np.ascontiguousarray(E, dtype=np.float64)
        out = np.empty_like(a)
        _check(lib().edus_rhs(self.h, _dp(e), _dp(a.view(np.float64)), _dp(out.view(np.float64))))
        return out

    def rhs_peierls(self, E, Av, rho) -> np.ndarray:
        a = _c128(rho).reshape(self._shape())
        e = np.ascontiguousarray(E, dtype=np.float64)
        av = np.ascontiguousarray(Av, dtype=np.float64)
        out = np.empty_like(a)
        _check(lib().edus_rhs_peierls(self.h, _dp(e), _dp(av), _dp(a.view(np.float64)), _dp(out.view(np.float64))))
        return out

    def observables_peierls(self, Av):
        pb, pw, v = np.empty(self.nb), np.empty(self.nb), np.empty(6)
        av = np.ascontiguousarray(Av, dtype=np.float64)
        _check(lib().edus_observables_peierls(self.h, _dp(av), _dp(pb), _dp(pw), _dp(v)))
        return pb, pw, v

    def rk4_steps(self, dt: float, E) -> None:
        e = np.ascontiguousarray(E, dtype=np.float64).reshape(-1, 3, 3)
        _check(lib().edus_rk4_steps(self.h, float(dt), e.shape[0], _dp(e)))

    def ab_steps(self, dt: float, order: int, E) -> None:
        e = np.ascontiguousarray(E, dtype=np.float64).reshape(-1, 3)
        _check(lib().edus_ab_steps(self.h, float(dt), int(order), e.shape[0], _dp(e)))

    def ab_reset(self):
        _check(lib().edus_ab_reset(self.h))

    def ab_seed(self, dt: float, order: int, E) -> None:
        """edus_ab_seed: order-1 RK4 steps (E [order-1][3][3]) that fill the AB history."""
        e = np.ascontiguousarray(E, dtype=np.float64).reshape(-1, 3, 3)
        _check(lib().edus_ab_seed(self.h, float(dt), int(order), _dp(e)))

    def observables(self):
        pb, pw, v = np.empty(self.nb), np.empty(self.nb), np.empty(6)
        _check(lib().edus_observables(self.h, _dp(pb), _dp(pw), _dp(v)))
        return pb, pw, v

    def check(self):
        t, h = C.c_double(), C.c_double()
        _check(lib().edus_check(self.h, C.byref(t), C.byref(h)))
        return t.value, h.value

    def monitor_limits(self, trace_limit: float, herm_limit: float) -> None:
        """edus_monitor_limits: step calls fail with EDUS_ERR_MONITOR beyond these (0: off)."""
        _check(lib().edus_monitor_limits(self.h, float(trace_limit), float(herm_limit)))

    def synchronize(self):
        _check(lib().edus_synchronize(self.h))

    def graph(self, steps_per_graph: int):
        """edus_graph: replay RK4 steps from hipGraphs of this many steps (0: off)."""
        _check(lib().edus_graph(self.h, int(steps_per_graph)))

    def timing_enable(self, on: bool = True):
        _check(lib().edus_timing_enable(self.h, int(bool(on))))

    def timing_get(self, cls: str):
        ms, n = C.c_double(), C.c_long()
        _check(lib().edus_timing_get(self.h, cls.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().edus_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, t: Tuning) -> None:
        _check(lib().edus_set_tuning(self.h, C.byref(t)))


def kgrad_stencil(A, N0: int, N1: int) -> np.ndarray:
    """edus_kgrad_stencil: the finite-difference stencil of gradient_space "k"
    for lattice A (columns: lattice vectors, bohr) and an N0 x N1 mesh, rows
    (shell, m0, m1, bx, by, bz, w).  Host only."""
    a = (C.c_double * 9)(*np.asarray(A, dtype=float).ravel())
    n = lib().edus_kgrad_stencil(C.byref(a), int(N0), int(N1), None, 0)
    if n < 0:
        raise EdusError(n, "no finite-difference stencil for this lattice and mesh")
    out = np.empty((n, 7))
    lib().edus_kgrad_stencil(C.byref(a), int(N0), int(N1), _dp(out), n)
    return out


def lindhard(U, eps, N0: int, N1: int, filledbands: int) -> np.ndarray:
    """edus_lindhard: the Lindhard response X(q) [q][a][b] of the screened-potential
    application (Screened_potential.cpp:17-58) for eigenvectors U [k][a][n]
    (bands in columns, ascending) and energies eps [k][n], computed on the device."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    nk, nb, nb2 = U.shape
    if nb != nb2 or nk != int(N0) * int(N1) or e.shape != (nk, nb):
        raise ValueError("U must be [N0*N1][nb][nb] and eps [N0*N1][nb]")
    X = np.empty_like(U)
    _check(lib().edus_lindhard(nb, int(N0), int(N1), int(filledbands), _dp(U.view(np.float64)), _dp(e),
                               _dp(X.view(np.float64))))
    return X


def lindhard_device(nb: int, N0: int, N1: int, filledbands: int, U_ptr: int, eps_ptr: int, work_ptr: int,
                    X_ptr: int, stream: int = 0):
    """edus_lindhard_device on device pointers (U [k][a][n], eps [k][n], work of
    Nk*nb*nb complex, X out as planes [nb*nb][Nk]) in the given HIP stream."""
    _check(lib().edus_lindhard_device(int(nb), int(N0), int(N1), int(filledbands), C.c_void_p(U_ptr),
                                      C.c_void_p(eps_ptr), C.c_void_p(work_ptr), C.c_void_p(X_ptr),
                                      C.c_void_p(stream)))


def lindhard_device_noalloc(nb: int, N0: int, N1: int, filledbands: int, U_ptr: int, eps_ptr: int, X_ptr: int,
                            stream: int = 0) -> None:
    """edus_lindhard_device_noalloc on device pointers (no workspace, no allocation)."""
    _check(lib().edus_lindhard_device_noalloc(int(nb), int(N0), int(N1), int(filledbands), C.c_void_p(U_ptr),
                                              C.c_void_p(eps_ptr), C.c_void_p(X_ptr), C.c_void_p(stream)))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """edus_mlp_*: the per-k network of the learned coarse propagator."""

    def __init__(self, din: int, dh: int, dout: int, seed: int = 0):
        self.din, self.dh, self.dout = din, dh, dout
        self.h = lib().edus_mlp_create(din, dh, dout, seed)
        if not self.h:
            raise EdusError(-1, lib().edus_last_error().decode())
        n = C.c_long()
        _check(lib().edus_mlp_num_params(self.h, C.byref(n)))
        self.nparam = n.value

    def close(self):
        if getattr(self, "h", None):
            lib().edus_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_params(self):
        p = np.empty(self.nparam, dtype=np.float32)
        _check(lib().edus_mlp_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.nparam
        _check(lib().edus_mlp_set_params(self.h, _fp(a)))

    def forward(self, X):
        x = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty((x.shape[0], self.dout), dtype=np.float32)
        _check(lib().edus_mlp_forward(self.h, x.shape[0], _fp(x), _fp(y)))
        return y

    def gradients(self, X, T):
        x = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = C.c_float()
        _check(lib().edus_mlp_gradients(self.h, x.shape[0], _fp(x), _fp(t), _fp(g), C.byref(loss)))
        return loss.value, g

    def train(self, X, T, epochs: int, batch: int, lr: float, seed: int = 0):
        x = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.empty(max(epochs, 1), dtype=np.float32)
        _check(lib().edus_mlp_train(self.h, x.shape[0], _fp(x), _fp(t), int(epochs), int(batch), float(lr),
                                    int(seed), _fp(hist)))
        return hist[:epochs]


class Batch:
    """edus_batch_*: many contexts advanced together.  The members (Engine
    objects) must stay alive as long as the batch."""

    def __init__(self, members):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        self.h = lib().edus_batch_create(len(self.members), arr)
        if not self.h:
            raise EdusError(-1, lib().edus_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            lib().edus_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def n(self):
        return len(self.members)

    def groups(self):
        ng = C.c_int()
        mg = (C.c_int * self.n)()
        _check(lib().edus_batch_groups(self.h, C.byref(ng), mg))
        return ng.value, list(mg)

    def graph(self, steps_per_graph: int):
        _check(lib().edus_batch_graph(self.h, int(steps_per_graph)))

    def rk4_steps(self, dt: float, E) -> None:
        """E[s][j][member][3] at t_s, t_s + dt/2, t_s + dt."""
        e = np.ascontiguousarray(E, dtype=np.float64).reshape(-1, 3, self.n, 3)
        _check(lib().edus_batch_rk4_steps(self.h, float(dt), e.shape[0], _dp(e)))

    def store(self, i: int):
        _check(lib().edus_batch_store(self.h, int(i)))

    def load(self, i: int):
        _check(lib().edus_batch_load(self.h, int(i)))

    def observables(self, i: int):
        nb = self.members[i].nb
        pb, pw, v = np.empty(nb), np.empty(nb), np.empty(6)
        _check(lib().edus_batch_observables(self.h, int(i), _dp(pb), _dp(pw), _dp(v)))
        return pb, pw, v

    def check(self):
        t, h = C.c_double(), C.c_double()
        _check(lib().edus_batch_check(self.h, C.byref(t), C.byref(h)))
        return t.value, h.value

    def synchronize(self):
        _check(lib().edus_batch_synchronize(self.h))


def _sim_params(p: dict, precision: int, layout: int) -> _SimParams:
    s = _SimParams()
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
    s.precision = int(precision)
    s.layout = int(layout)
    s.ab_seed = int(bool(p.get("ab_seed", False)))
    s.peierls = int(bool(p.get("peierls", False)))
    s.gradient_space = 1 if p.get("gradient_space", "R") == "k" else 0
    s.write_h5 = int(bool(p.get("write_h5", False)))
    toprint = p.get("toprint", {})
    s.toprint_dmk_wannier = int(str(toprint.get("DMk_wannier", "false")) == "true")
    s.toprint_dmk_bloch = int(str(toprint.get("DMk_bloch", "false")) == "true")
    return s


class Simulation:
    """Reference-format simulation (edus_sim_*): host set-up + engine."""

    def __init__(self, params: dict, precision: int = PREC_FP64, layout: int = LAYOUT_FULL):
        self._p = _sim_params(params, precision, layout)
        self.h = lib().edus_sim_create(C.byref(self._p))
        if not self.h:
            raise EdusError(-1, lib().edus_last_error().decode())
        nb, n0, n1, ns = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        dt = C.c_double()
        _check(lib().edus_sim_info(self.h, C.byref(nb), C.byref(n0), C.byref(n1), C.byref(ns), C.byref(dt)))
        self.nb, self.N0, self.N1, self.nsteps, self.dt = nb.value, n0.value, n1.value, ns.value, dt.value
        self.Nk = self.N0 * self.N1
        self.engine = Engine(self.nb, self.N0, self.N1, None, _handle=lib().edus_sim_engine(self.h), _owner=self)

    def close(self):
        if getattr(self, "h", None):
            lib().edus_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def time(self) -> float:
        return lib().edus_sim_time(self.h)

    def field(self, t: float) -> np.ndarray:
        out = np.empty(3)
        _check(lib().edus_sim_field(self.h, float(t), _dp(out)))
        return out

    def get(self, name: str) -> np.ndarray:
        n = lib().edus_sim_get(self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        if name == "eps":
            out = np.empty(n)
            lib().edus_sim_get(self.h, name.encode(), _dp(out))
            return out.reshape(self.Nk, self.nb)
        out = np.empty(n, dtype=np.complex128)
        if lib().edus_sim_get(self.h, name.encode(), _dp(out.view(np.float64))) < 0:
            raise EdusError(-1, lib().edus_last_error().decode())
        if name == "hartree":
            return out.reshape(self.nb, self.nb)
        return out.reshape(-1, self.nb, self.nb)

    def steps(self, n: int):
        _check(lib().edus_sim_steps(self.h, int(n)))

    def write_outputs(self, directory: str, nsteps: int | None = None) -> int:
        """edus_sim_write_outputs: propagate and write the reference's text files."""
        k = C.c_int()
        _check(lib().edus_sim_write_outputs(self.h, directory.encode(), -1 if nsteps is None else int(nsteps),
                                            C.byref(k)))
        return k.value

    def screened_potential(self) -> dict:
        """edus_sim_screened (ScreenedPotential::Calculate): the R-space fields
        X, V, Eps, InvEps, W, each [R][nb][nb] over the FFT grid."""
        names = ("X", "V", "Eps", "InvEps", "W")
        out = {k: np.empty((self.Nk, self.nb, self.nb), dtype=np.complex128) for k in names}
        _check(lib().edus_sim_screened(self.h, *(_dp(out[k].view(np.float64)) for k in names)))
        return out

    def write_screened(self, directory: str):
        """edus_sim_write_screened: X0.txt, V0.txt, Epsilon0.txt, InverseEpsilon0.txt, W0.txt."""
        _check(lib().edus_sim_write_screened(self.h, directory.encode()))

    def propagate(self, nsteps: int | None = None) -> dict:
        n = self.nsteps if nsteps is None else int(nsteps)
        times, laser = np.empty(n), np.empty((n, 3))
        pb, pw, vel = np.empty((n, self.nb)), np.empty((n, self.nb)), np.empty((n, 6))
        k = C.c_int()
        _check(lib().edus_sim_propagate(self.h, n, _dp(times), _dp(laser), _dp(pb), _dp(pw), _dp(vel),
                                        C.byref(k)))
        k = k.value
        return {"time": times[:k], "laser": laser[:k], "pop_bloch": pb[:k], "pop_wannier": pw[:k],
                "velocity": vel[:k]}
