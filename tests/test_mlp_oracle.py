"""CPU: the MLP oracle (oracle/mlp_ref.py) pinned before it judges the GPU.

- the float64 backpropagation against central finite differences (every
  parameter block, so a dropped term or transposed factor fails);
- BF16 rounding against hand-worked bit patterns (ties to even, carries into
  the exponent);
- the rounded-arithmetic gradients close to the exact ones at the size of
  BF16 rounding.
"""
import numpy as np

import mlp_ref

DIN, DH, DOUT = 5, 7, 3


def _params(seed):
    rng = np.random.default_rng(seed)
    n = DIN * DH + DH + DH * DH + DH + DH * DOUT + DOUT
    return rng.standard_normal(n) * 0.4


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    X = rng.standard_normal((11, DIN))
    T = rng.standard_normal((11, DOUT))
    g = mlp_ref.grads_exact(p, X, T, DIN, DH, DOUT)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp_ref.loss_exact(p + e, X, T, DIN, DH, DOUT) - mlp_ref.loss_exact(p - e, X, T, DIN, DH, DOUT)) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-8 * max(1.0, np.abs(fd).max())


def test_bf16_rounding_known_values():
    # (float32 bits, expected BF16 bits), worked by hand:
    cases = [(0x3F800000, 0x3F80),  # 1.0 exact
             (0x3F808000, 0x3F80),  # 1 + 2^-8: tie, kept LSB even -> down
             (0x3F818000, 0x3F82),  # 1 + 2^-7 + 2^-8: tie, kept LSB odd -> up
             (0x3F80C000, 0x3F81),  # above the tie -> up
             (0x7F61B1E6, 0x7F62),  # 3.0e38 -> up
             (0xC0200000, 0xC020),  # -2.5 exact
             (0x477FFF00, 0x4780)]  # 65535 -> 65536, carry into the exponent
    x = np.array([c[0] for c in cases], dtype=np.uint32).view(np.float32)
    got = mlp_ref.bf16(x).view(np.uint32) >> 16
    assert got.tolist() == [c[1] for c in cases]


def test_rounded_gradients_close_to_exact():
    rng = np.random.default_rng(3)
    p = _params(4).astype(np.float32)
    X = rng.standard_normal((64, DIN)).astype(np.float32)
    T = rng.standard_normal((64, DOUT)).astype(np.float32)
    _, g = mlp_ref.grads(p, X, T, DIN, DH, DOUT)
    # the exact gradient of the same network evaluated at BF16-rounded inputs and weights
    pb = p.copy()
    W1, b1, W2, b2, W3, b3 = mlp_ref.split(pb, DIN, DH, DOUT)
    for W in (W1, W2, W3):
        W[...] = mlp_ref.bf16(W)
    ge = mlp_ref.grads_exact(pb, mlp_ref.bf16(X).astype(np.float64), T, DIN, DH, DOUT)
    assert np.linalg.norm(g - ge) <= 2e-2 * np.linalg.norm(ge)
