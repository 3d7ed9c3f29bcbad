"""GPU: the per-k MLP kernels (BF16 matrix cores, FP32 accumulation) against
the CPU oracle oracle/mlp_ref.py (same roundings), and training.

Bars: forward outputs within 2e-3 of the output scale (a BF16 rounding of a
hidden activation may land on the other side where the FP32 sums differ in
order); every parameter block's gradient within 1e-2 relative (Frobenius).
"""
import sys

import numpy as np
import pytest

import edus

mlp_ref = pytest.importorskip("mlp_ref")

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("din,dh,dout,rows", [(84, 256, 64, 3000), (20, 64, 10, 777), (5, 16, 3, 33)])
def test_forward_and_gradients_match_oracle(din, dh, dout, rows):
    rng = np.random.default_rng(din + rows)
    net = edus.Mlp(din, dh, dout, seed=5)
    p = net.get_params()
    # a trained-looking parameter set: random biases too
    p = p + 0.05 * rng.standard_normal(p.size).astype(np.float32)
    net.set_params(p)
    X = rng.standard_normal((rows, din)).astype(np.float32)
    T = rng.standard_normal((rows, dout)).astype(np.float32)
    y = net.forward(X)
    y_ref, _ = mlp_ref.forward(p, X, din, dh, dout)
    ferr = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    loss, g = net.gradients(X, T)
    loss_ref, g_ref = mlp_ref.grads(p, X, T, din, dh, dout)
    _say("shape", din, dh, dout, rows, "forward rel err", ferr, "loss", loss, loss_ref)
    assert ferr <= 2e-3
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    blocks = mlp_ref.split(np.arange(p.size), din, dh, dout)
    for name, idx in zip(("W1", "b1", "W2", "b2", "W3", "b3"), blocks):
        i = idx.ravel()
        rel = np.linalg.norm(g[i] - g_ref[i]) / max(np.linalg.norm(g_ref[i]), 1e-30)
        _say("  grad", name, "rel err", rel)
        assert rel <= 1e-2, name


def test_training_fits_a_smooth_map():
    """Adam on the GPU fits a smooth 8 -> 4 map; the loss drops by 10x."""
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(8192, 8)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0] + X[:, 1]), X[:, 2] * X[:, 3], np.cos(2 * X[:, 4]), X[:, 5] - X[:, 6] ** 2],
                 axis=1).astype(np.float32)
    net = edus.Mlp(8, 64, 4, seed=1)
    hist = net.train(X, T, epochs=30, batch=256, lr=3e-3, seed=2)
    _say("loss per epoch", hist[0], hist[-1])
    assert hist[-1] < 0.1 * hist[0]
    y = net.forward(X[:512])
    assert np.sqrt(np.mean((y - T[:512]) ** 2)) < 0.15
