"""Planner surrogate network: the NumPy oracle pinned by finite differences and closed forms
(CPU), and the HIP forward / backward / Adam kernels against the oracle (GPU).

Tolerances (FP32 kernels vs float64 oracle): forward and gradients within 1e-4 relative to
the largest magnitude of the compared array; Adam parameters within 5e-3 * lr per step."""
from __future__ import annotations

import numpy as np
import pytest

import es_mlp
import es_mlp_oracle as MO


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 64))
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2], np.tanh(X[:, 3:6].sum(1))], 1)
    return X, T


def test_pack_roundtrip():
    rng = np.random.default_rng(0)
    parts = [rng.normal(size=s).astype(np.float32)
             for s in ((64, 128), (128,), (128, 128), (128,), (128, 2), (2,))]
    v = es_mlp.pack(*parts)
    assert v.size == es_mlp.N_PARAMS == 26896
    for a, b in zip(parts, es_mlp.unpack(v)):
        assert np.array_equal(a, b)


def test_oracle_gradient_is_finite_difference_of_loss():
    rng = np.random.default_rng(1)
    v = es_mlp.init_params(3).astype(np.float64)
    v += rng.normal(scale=0.05, size=v.size)        # nonzero biases and padded columns too
    X, T = _data(9, 2)
    w = (1.0, 0.7)
    L, g = MO.gradients(v, X, T, w)
    assert abs(L - MO.loss(v, X, T, w)) < 1e-12
    sizes = [64 * 128, 128, 128 * 128, 128, 128 * 16, 16]
    starts = np.cumsum([0] + sizes[:-1])
    picks = []
    for s, n in zip(starts, sizes):                   # coordinates from every block
        picks += list(s + rng.choice(n, size=min(n, 40), replace=False))
    h = 1e-6
    for i in picks:
        e = np.zeros_like(v)
        e[i] = h
        fd = (MO.loss(v + e, X, T, w) - MO.loss(v - e, X, T, w)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i
    # padded output columns 2..15 carry no gradient
    W3 = g[starts[4]:starts[4] + sizes[4]].reshape(128, 16)
    assert np.all(W3[:, 2:] == 0) and np.all(g[starts[5] + 2:] == 0)


def test_oracle_forward_closed_forms():
    X, _ = _data(7, 4)
    z = np.zeros(es_mlp.N_PARAMS)
    b3 = np.array([0.25, -1.5])
    v = es_mlp.pack(np.zeros((64, 128)), np.zeros(128), np.zeros((128, 128)), np.zeros(128),
                    np.zeros((128, 2)), b3)
    Y, _ = MO.forward(v, X)
    assert np.allclose(Y, b3, atol=1e-7)              # zero weights: the output bias
    W1 = np.zeros((64, 128)); W1[0, 0] = 1.0; W1[1, 1] = -1.0
    W3 = np.zeros((128, 2)); W3[0, 0] = 1.0; W3[1, 1] = 2.0
    v = es_mlp.pack(W1, np.zeros(128), np.eye(128), np.zeros(128), W3, np.zeros(2))
    Y, _ = MO.forward(v, X)
    assert np.allclose(Y[:, 0], np.maximum(X[:, 0], 0), atol=1e-6)
    assert np.allclose(Y[:, 1], 2 * np.maximum(-X[:, 1], 0), atol=1e-6)
    assert z.size == v.size


def test_oracle_adam_first_step_is_signed_lr():
    g = np.array([3.0, -2e-3, 0.0, 5e-1])
    v1, m, s = MO.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    assert np.allclose(v1[[0, 1, 3]], -0.01 * np.sign(g[[0, 1, 3]]), rtol=1e-5)
    assert v1[2] == 0.0


@pytest.mark.gpu
def test_gpu_forward_matches_oracle():
    net = es_mlp.Mlp(batch_cap=512, seed=5)
    X, _ = _data(1000, 6)                            # two chunks, ragged last one
    Y = net.forward(X)
    want, _ = MO.forward(net.params(), X)
    assert np.max(np.abs(Y - want)) <= 1e-4 * max(1.0, np.max(np.abs(want)))


@pytest.mark.gpu
def test_gpu_gradients_match_oracle():
    rng = np.random.default_rng(7)
    v = es_mlp.init_params(8) + rng.normal(scale=0.02, size=es_mlp.N_PARAMS).astype(np.float32)
    net = es_mlp.Mlp(batch_cap=256, params=v)
    v = net.params()                                  # padded columns zeroed by the library
    for n in (16, 200, 256):
        X, T = _data(n, n)
        L, g = net.gradients(X, T, (1.0, 0.5))
        Lw, gw = MO.gradients(v, X, T, (1.0, 0.5))
        assert abs(L - Lw) <= 1e-4 * max(1.0, abs(Lw))
        assert np.max(np.abs(g - gw)) <= 1e-4 * np.max(np.abs(gw))


@pytest.mark.gpu
def test_gpu_adam_steps_match_oracle():
    net = es_mlp.Mlp(batch_cap=128, seed=9)
    v = net.params().astype(np.float64)
    m = np.zeros_like(v)
    s = np.zeros_like(v)
    lr = 1e-3
    for t in range(1, 4):
        X, T = _data(100, 10 + t)
        net.adam_step(X, T, lr=lr)
        _, g = MO.gradients(v, X, T)
        v, m, s = MO.adam(v, g, m, s, t, lr=lr)
        got = net.params()
        assert np.max(np.abs(got - v)) <= 5e-3 * lr * t


@pytest.mark.gpu
def test_gpu_training_fits_smooth_target():
    X, T = _data(8192, 21)
    net = es_mlp.Mlp(batch_cap=512, seed=22)
    L0 = MO.loss(net.params(), X, T)
    last, ms = net.fit(X, T, epochs=40, batch=256, lr=2e-3)
    L1 = MO.loss(net.params(), X, T)
    print(f"mlp fit: loss {L0:.4f} -> {L1:.4f} in {ms:.1f} ms", flush=True)
    assert L1 < 0.1 * L0 and np.isfinite(last)
