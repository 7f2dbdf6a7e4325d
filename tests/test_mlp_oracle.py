"""The warm-start network's CPU oracle (oracle/mlp.py) pinned by central
finite differences of its own loss, on a small network with every layer
kind (input, two tanh hidden layers, linear output)."""
import numpy as np

from oracle.mlp import forward, layer_slices, loss_grad


def test_gradient_matches_finite_differences():
    widths = [2, 7, 5, 3]
    _, n = layer_slices(widths)
    rng = np.random.default_rng(3)
    p = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((6, 2))
    t = rng.standard_normal((6, 3))
    loss, g = loss_grad(p, widths, x, t)
    h = 1e-6
    fd = np.empty(n)
    for k in range(n):
        e = np.zeros(n)
        e[k] = h
        fd[k] = (loss_grad(p + e, widths, x, t)[0] - loss_grad(p - e, widths, x, t)[0]) / (2 * h)
    assert np.abs(fd - g).max() <= 1e-8 * max(1.0, np.abs(g).max())
    assert loss == np.mean((forward(p, widths, x)[-1] - t) ** 2)


def test_linear_network_closed_form():
    """One linear layer: y = x W^T + b, dL/dW = 2/n (y - t)^T x."""
    widths = [3, 2]
    rng = np.random.default_rng(4)
    W, b = rng.standard_normal((2, 3)), rng.standard_normal(2)
    p = np.concatenate([W.reshape(-1), b])
    x, t = rng.standard_normal((5, 3)), rng.standard_normal((5, 2))
    y = x @ W.T + b
    loss, g = loss_grad(p, widths, x, t)
    assert np.allclose(g[:6], (2 / 10) * ((y - t).T @ x).reshape(-1), rtol=1e-14, atol=1e-15)
    assert np.allclose(g[6:], (2 / 10) * (y - t).sum(0), rtol=1e-14, atol=1e-15)
    assert abs(loss - np.mean((y - t) ** 2)) <= 1e-15
