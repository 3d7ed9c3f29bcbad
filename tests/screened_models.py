"""Models with a closed-form Lindhard response (used by the oracle and GPU
tests of the screened-potential application).

Gauge-phased two-band blocks.  Block b occupies orbitals (2b, 2b+1); its
valence band is column b, its conduction band column nblocks + b:

    U_k[2b, b]   = cos t_b             U_k[2b, nblocks+b]   = -sin t_b
    U_k[2b+1, b] = e^{i phi_b(k)} sin t_b   U_k[2b+1, nblocks+b] = e^{i phi_b(k)} cos t_b

with a non-linear periodic gauge phase phi_b(k) and dispersive energies
e_v,b(k) < e_c,b(k).  In the response
    X_ab(q) = 1/Nk sum_k sum_{v,c} U_k(a,v) U_k+q(b,c) conj(U_k+q(a,c)) conj(U_k(b,v))
              / (e_v(k) - e_c(k+q))
only v and c of the block holding both a and b contribute, and by hand:

    X_{2b,2b}(q) = X_{2b+1,2b+1}(q) = c^2 s^2 <1 / d_b(k, q)>_k
    X_{2b,2b+1}(q) = -c^2 s^2 <exp(i (phi_b(k+q) - phi_b(k))) / d_b(k, q)>_k
    X_{2b+1,2b}(q) = conj X_{2b,2b+1}(q),      d_b(k, q) = e_v,b(k) - e_c,b(k+q),

all other elements zero.  The phase factor fixes the direction of the shift
(k + q, not k - q), the order of the four eigenvector factors and the sign of
the denominator; the dispersion of both bands fixes which energy is taken at
k and which at k + q.
"""
import numpy as np


def _grid(N0, N1):
    i0, i1 = np.meshgrid(np.arange(N0), np.arange(N1), indexing="ij")
    return (2 * np.pi * i0 / N0).ravel(), (2 * np.pi * i1 / N1).ravel()


def block_params(nblocks):
    b = np.arange(nblocks)
    return {
        "theta": 0.3 + 0.37 * b,
        "amp0": 0.9 - 0.11 * b,
        "amp1": 0.5 + 0.07 * b,
        "shift": 0.4 * b + 0.2,
    }


def gauge_block_model(nblocks, N0, N1):
    """U [Nk][nb][nb], eps [Nk][nb] and filledbands (= nblocks)."""
    x0, x1 = _grid(N0, N1)
    nb, nk = 2 * nblocks, N0 * N1
    p = block_params(nblocks)
    U = np.zeros((nk, nb, nb), dtype=np.complex128)
    eps = np.zeros((nk, nb))
    for b in range(nblocks):
        c, s = np.cos(p["theta"][b]), np.sin(p["theta"][b])
        ph = np.exp(1j * phase(b, x0, x1, p))
        U[:, 2 * b, b] = c
        U[:, 2 * b + 1, b] = ph * s
        U[:, 2 * b, nblocks + b] = -s
        U[:, 2 * b + 1, nblocks + b] = ph * c
        ev, ec = energies(b, nblocks, x0, x1)
        eps[:, b] = ev
        eps[:, nblocks + b] = ec
    return U, eps, nblocks


def phase(b, x0, x1, p):
    return p["amp0"][b] * np.sin(x0 + p["shift"][b]) + p["amp1"][b] * np.cos(2 * x1) + 0.1 * np.sin(x0 + x1)


def energies(b, nblocks, x0, x1):
    ev = -1.5 + 0.3 * b / max(1, nblocks) + 0.2 * np.cos(x0 + b) + 0.1 * np.sin(x1)
    ec = 1.0 + 0.1 * b + 0.25 * np.sin(x1 + b) + 0.1 * np.cos(2 * x0)
    return ev, ec


def _shifted(a, q0):
    """w[k0, q1, k1] = a[(k0 + q0) % N0, (k1 + q1) % N1] for a grid array a (a view)."""
    r = np.roll(a, -q0, axis=0)
    ext = np.concatenate([r, r], axis=1)
    return np.lib.stride_tricks.sliding_window_view(ext, a.shape[1], axis=1)[:, : a.shape[1], :]


def closed_form(nblocks, N0, N1):
    """X [Nk][nb][nb] of gauge_block_model from the formulas above."""
    x0, x1 = _grid(N0, N1)
    nb, nk = 2 * nblocks, N0 * N1
    p = block_params(nblocks)
    X = np.zeros((N0, N1, nb, nb), dtype=np.complex128)
    for b in range(nblocks):
        c2s2 = (np.cos(p["theta"][b]) * np.sin(p["theta"][b])) ** 2
        eph = np.exp(1j * phase(b, x0, x1, p)).reshape(N0, N1)
        ev, ec = (e.reshape(N0, N1) for e in energies(b, nblocks, x0, x1))
        for q0 in range(N0):
            inv_d = 1.0 / (ev[:, None, :] - _shifted(ec, q0))  # [k0, q1, k1]
            diag = c2s2 * inv_d.sum(axis=(0, 2)) / nk
            off = -c2s2 * np.einsum("iqj,iqj,ij->q", _shifted(eph, q0), inv_d, np.conj(eph)) / nk
            X[q0, :, 2 * b, 2 * b] = diag
            X[q0, :, 2 * b + 1, 2 * b + 1] = diag
            X[q0, :, 2 * b, 2 * b + 1] = off
            X[q0, :, 2 * b + 1, 2 * b] = np.conj(off)
    return X.reshape(nk, nb, nb)


def random_bands(nb, N0, N1, seed):
    """Eigenvectors and ascending energies of random Hermitian H(k)."""
    rng = np.random.default_rng(seed)
    nk = N0 * N1
    a = rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))
    H = 0.5 * (a + a.conj().transpose(0, 2, 1)) + np.diag(np.linspace(-3.0, 3.0, nb))[None]
    w, U = np.linalg.eigh(H)
    return U, w
