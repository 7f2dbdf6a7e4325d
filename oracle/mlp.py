"""CPU oracle of the warm-start network (test infrastructure only: the tests
use it to check the library's device kernels; nothing in the product path
imports it).

FP64 numpy restatement of the network the library trains on the device
(qd_imsrg_amd/csrc/mlp.hip): layer l maps A (batch x in) to
Z = A W^T + b, tanh on hidden layers, linear output; parameters are one flat
vector, per layer W (out x in, row-major) then b; loss = mean over batch and
outputs of (Y - T)^2.  No counterpart in the reference.

Pinned by central finite differences of the loss (tests/test_mlp_oracle.py),
which a wrong transpose, a missing tanh derivative or a dropped bias
gradient fails.
"""
from __future__ import annotations

import numpy as np


def layer_slices(widths):
    """(W slice, b slice, in, out) per layer of the flat parameter vector."""
    out, k = [], 0
    for a, b in zip(widths[:-1], widths[1:]):
        ws = slice(k, k + a * b)
        k += a * b
        bs = slice(k, k + b)
        k += b
        out.append((ws, bs, a, b))
    return out, k


def forward(p, widths, x):
    """Activations of every layer (input first)."""
    acts = [np.asarray(x, dtype=np.float64)]
    sl, _ = layer_slices(widths)
    for i, (ws, bs, a, b) in enumerate(sl):
        W = p[ws].reshape(b, a)
        z = acts[-1] @ W.T + p[bs]
        acts.append(np.tanh(z) if i + 1 < len(sl) else z)
    return acts


def loss_grad(p, widths, x, t):
    """Mean squared error and its gradient with respect to every parameter."""
    p = np.asarray(p, dtype=np.float64)
    acts = forward(p, widths, x)
    y = acts[-1]
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    g = np.zeros_like(p)
    sl, _ = layer_slices(widths)
    G = 2.0 * (y - t) / n
    for i in range(len(sl) - 1, -1, -1):
        ws, bs, a, b = sl[i]
        g[ws] = (G.T @ acts[i]).reshape(-1)
        g[bs] = G.sum(axis=0)
        if i > 0:
            W = p[ws].reshape(b, a)
            G = (G @ W) * (1.0 - acts[i] ** 2)
    return loss, g
