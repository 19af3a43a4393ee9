"""CPU ORACLE for the planner surrogate network.  TEST INFRASTRUCTURE ONLY: imported by tests/
and bench.py as the checker, never by the engine.

Float64 NumPy restatement of the 64-128-128-2 ReLU MLP, its weighted mean-squared-error loss,
the loss gradient by backpropagation, and the bias-corrected Adam update, on the same packed
parameter layout as include/es_mlp.h.  There is no reference counterpart (the reference has
no learned planner); the gradient is pinned in tests by central finite differences of the
loss, and the forward pass by closed-form special cases.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, OUT_PAD = 64, 128, 2, 16


def unpack(v):
    v = np.asarray(v, dtype=np.float64)
    o, out = 0, []
    for shape in ((IN, HID), (HID,), (HID, HID), (HID,), (HID, OUT_PAD), (OUT_PAD,)):
        n = int(np.prod(shape))
        out.append(v[o:o + n].reshape(shape))
        o += n
    return out


def forward(v, X):
    W1, b1, W2, b2, W3, b3 = unpack(v)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1 + b1
    H1 = np.maximum(Z1, 0)
    Z2 = H1 @ W2 + b2
    H2 = np.maximum(Z2, 0)
    Y = H2 @ W3 + b3
    return Y[:, :OUT], (X, Z1, H1, Z2, H2, Y)


def loss(v, X, T, w=(1.0, 1.0)):
    Y, _ = forward(v, X)
    d = Y - np.asarray(T, dtype=np.float64)
    return float(np.mean(np.sum(np.asarray(w) * d * d, axis=1)))


def gradients(v, X, T, w=(1.0, 1.0)):
    """(loss, gradient packed like v) by backpropagation."""
    W1, b1, W2, b2, W3, b3 = unpack(v)
    _, (X, Z1, H1, Z2, H2, Y) = forward(v, X)
    n = X.shape[0]
    T = np.asarray(T, dtype=np.float64)
    dY = np.zeros_like(Y)
    d = Y[:, :OUT] - T
    dY[:, :OUT] = 2.0 * np.asarray(w) * d / n
    gW3 = H2.T @ dY
    gb3 = dY.sum(0)
    dH2 = (dY @ W3.T) * (Z2 > 0)
    gW2 = H1.T @ dH2
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ W2.T) * (Z1 > 0)
    gW1 = X.T @ dH1
    gb1 = dH1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(np.mean(np.sum(np.asarray(w) * d * d, axis=1))), g


def adam(v, g, m, s, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step at step number t (1-based); returns (v, m, s)."""
    m = b1 * m + (1 - b1) * g
    s = b2 * s + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    sh = s / (1 - b2 ** t)
    return v - lr * mh / (np.sqrt(sh) + eps), m, s
