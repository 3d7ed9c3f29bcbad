"""CPU oracle of the per-k MLP (test infrastructure only: imported by tests/).

Same network and arithmetic as the engine's HIP kernels
(edus_2.0_amd/csrc/learn/mlp.hip): three dense layers din -> dh -> dh -> dout,
tanh after the first two, inputs, weights and hidden activations rounded to
BF16 (round to nearest even), products accumulated in FP32, FP32 biases.
Loss L = mean over rows and outputs of (Y - T)^2.  The backward pass keeps
the kernels' roundings: dY and the hidden-layer gradients are BF16 before
they enter a product.

`grads_exact` is the same network in float64 without any rounding; it is
pinned against central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x007FFFFF) != 0
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.view(np.float32).reshape(a.shape)


def split(params, din, dh, dout):
    """Parameter block -> (W1, b1, W2, b2, W3, b3)."""
    p = np.asarray(params)
    shapes = [(din, dh), (dh,), (dh, dh), (dh,), (dh, dout), (dout,)]
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s))
        o += n
    assert o == p.size
    return out


def forward(params, X, din, dh, dout):
    W1, b1, W2, b2, W3, b3 = [np.asarray(a, dtype=np.float32) for a in split(params, din, dh, dout)]
    Xb = bf16(X)
    H1 = bf16(np.tanh(Xb @ bf16(W1) + b1))
    H2 = bf16(np.tanh(H1 @ bf16(W2) + b2))
    Y = H2 @ bf16(W3) + b3
    return Y.astype(np.float32), (Xb, H1, H2)


def grads(params, X, T, din, dh, dout):
    """Loss and parameter gradients with the kernels' roundings."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(a, dtype=np.float32) for a in split(params, din, dh, dout)]
    Y, (Xb, H1, H2) = forward(params, X, din, dh, dout)
    d = Y - np.asarray(T, dtype=np.float32)
    n = d.size
    loss = float(np.mean(d.astype(np.float64) ** 2))
    dY = bf16(2.0 * d / np.float32(n))
    gW3, gb3 = H2.T @ dY, dY.sum(0)
    dZ2 = bf16((dY @ bf16(W3).T) * (1.0 - H2 * H2))
    gW2, gb2 = H1.T @ dZ2, dZ2.sum(0)
    dZ1 = bf16((dZ2 @ bf16(W2).T) * (1.0 - H1 * H1))
    gW1, gb1 = Xb.T @ dZ1, dZ1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)]).astype(np.float32)
    return loss, g


def loss_exact(params, X, T, din, dh, dout):
    W1, b1, W2, b2, W3, b3 = split(np.asarray(params, dtype=np.float64), din, dh, dout)
    H1 = np.tanh(X @ W1 + b1)
    H2 = np.tanh(H1 @ W2 + b2)
    Y = H2 @ W3 + b3
    return float(np.mean((Y - T) ** 2))


def grads_exact(params, X, T, din, dh, dout):
    """float64 backpropagation without rounding."""
    W1, b1, W2, b2, W3, b3 = split(np.asarray(params, dtype=np.float64), din, dh, dout)
    X = np.asarray(X, dtype=np.float64)
    H1 = np.tanh(X @ W1 + b1)
    H2 = np.tanh(H1 @ W2 + b2)
    Y = H2 @ W3 + b3
    dY = 2.0 * (Y - T) / Y.size
    gW3, gb3 = H2.T @ dY, dY.sum(0)
    dZ2 = (dY @ W3.T) * (1.0 - H2 * H2)
    gW2, gb2 = H1.T @ dZ2, dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (1.0 - H1 * H1)
    gW1, gb1 = X.T @ dZ1, dZ1.sum(0)
    return np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
