"""Host side of the planner surrogate network (es_mlp_* in libes_sieve.so).

A 64-128-128-2 FP32 MLP with ReLU hidden layers whose forward pass, gradients and Adam updates
run on the GPU (hand-written MFMA GEMM tiles, es_mlp.hip).  Parameters are one packed vector
(W1 [64x128], b1, W2 [128x128], b2, W3 [128x16], b3 [16]; output columns 2..15 stay zero);
`pack`/`unpack` convert to and from per-layer arrays with W3 as [128x2].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import es_native as nat

IN, HID, OUT, OUT_PAD = 64, 128, 2, 16
N_PARAMS = IN * HID + HID + HID * HID + HID + HID * OUT_PAD + OUT_PAD


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _p(a):
    return a.ctypes.data_as(nat.f32p)


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    W3p = np.zeros((HID, OUT_PAD), dtype=np.float32)
    W3p[:, :OUT] = W3
    b3p = np.zeros(OUT_PAD, dtype=np.float32)
    b3p[:OUT] = b3
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3p), b3p]).astype(np.float32)


def unpack(v):
    v = np.asarray(v)
    o = 0
    out = []
    for shape in ((IN, HID), (HID,), (HID, HID), (HID,), (HID, OUT_PAD), (OUT_PAD,)):
        n = int(np.prod(shape))
        out.append(v[o:o + n].reshape(shape))
        o += n
    W1, b1, W2, b2, W3, b3 = out
    return W1, b1, W2, b2, W3[:, :OUT], b3[:OUT]


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform hidden weights, small output weights, zero biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (IN, HID)) * np.sqrt(6.0 / IN)
    W2 = rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6.0 / HID)
    W3 = rng.uniform(-1, 1, (HID, OUT)) * np.sqrt(1.0 / HID)
    return pack(W1, np.zeros(HID), W2, np.zeros(HID), W3, np.zeros(OUT))


class Mlp:
    """The network on the current GPU."""

    def __init__(self, batch_cap: int = 4096, params=None, seed: int = 0):
        h = C.c_void_p()
        nat.check(nat.lib().es_mlp_create(int(batch_cap), C.byref(h)))
        self._h = h
        self.batch_cap = batch_cap
        self.set_params(init_params(seed) if params is None else params)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                nat.lib().es_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_params(self, v) -> None:
        v = _f32(v)
        assert v.size == N_PARAMS
        nat.check(nat.lib().es_mlp_set_params(self._h, _p(v), N_PARAMS))

    def params(self) -> np.ndarray:
        v = np.zeros(N_PARAMS, dtype=np.float32)
        nat.check(nat.lib().es_mlp_get_params(self._h, _p(v), N_PARAMS))
        return v

    def forward(self, X) -> np.ndarray:
        X = _f32(X).reshape(-1, IN)
        Y = np.zeros((X.shape[0], OUT), dtype=np.float32)
        nat.check(nat.lib().es_mlp_forward(self._h, _p(X), X.shape[0], _p(Y)))
        return Y

    def gradients(self, X, T, w=(1.0, 1.0)):
        X, T, w = _f32(X).reshape(-1, IN), _f32(T).reshape(-1, OUT), _f32(w)
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = C.c_float(0)
        nat.check(nat.lib().es_mlp_gradients(self._h, _p(X), _p(T), X.shape[0], _p(w), _p(g),
                                              C.byref(loss)))
        return loss.value, g

    def adam_step(self, X, T, w=(1.0, 1.0), lr=1e-3, b1=0.9, b2=0.999, eps=1e-8) -> float:
        X, T, w = _f32(X).reshape(-1, IN), _f32(T).reshape(-1, OUT), _f32(w)
        loss = C.c_float(0)
        nat.check(nat.lib().es_mlp_adam_step(self._h, _p(X), _p(T), X.shape[0], _p(w), lr, b1,
                                              b2, eps, C.byref(loss)))
        return loss.value

    def fit(self, X, T, epochs: int, batch: int = 256, seed: int = 1, w=(1.0, 1.0), lr=1e-3,
            b1=0.9, b2=0.999, eps=1e-8):
        """Minibatch Adam over a device-resident training set; returns (last loss, device ms)."""
        X, T, w = _f32(X).reshape(-1, IN), _f32(T).reshape(-1, OUT), _f32(w)
        nat.check(nat.lib().es_mlp_upload_dataset(self._h, _p(X), _p(T), X.shape[0]))
        loss = C.c_float(0)
        ms = C.c_double(0)
        nat.check(nat.lib().es_mlp_train(self._h, epochs, batch, seed, _p(w), lr, b1, b2, eps,
                                         C.byref(loss), C.byref(ms)))
        return loss.value, ms.value
