"""Learned band predictor over the C ABI (gp_mlp_*): GPU features, BF16-MFMA
inference and training, and learned-band plans.

The network predicts y = log(d / e) (d: PTP distance, e: Euclidean distance
to the nearest source); the band schedule orders vertices by dhat = e*exp(y).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import GPError, check, lib, ptr

NPARAM = L.MLP_NPARAM
NFEAT = L.MLP_NFEAT


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).reshape(-1))


def features(mesh, sources):
    """(F float32[V, 10], e float32[V]) computed on the GPU."""
    src = _u32(sources)
    F = np.zeros((mesh.n_vertices(), NFEAT), np.float32)
    e = np.zeros(mesh.n_vertices(), np.float32)
    check(lib().gp_mlp_features(mesh.handle, ptr(src), len(src), ptr(F), ptr(e)))
    return F, e


class MLP:
    def __init__(self, seed: int = 0, params=None):
        h = C.c_void_p()
        check(lib().gp_mlp_create(seed, C.byref(h)))
        self._h = h
        if params is not None:
            self.set_params(params)

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.zeros(NPARAM, np.float32)
        check(lib().gp_mlp_params(self._h, ptr(out), None, NPARAM))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != NPARAM:
            raise GPError(L.ERR_CAPACITY, f"expected {NPARAM} parameters")
        check(lib().gp_mlp_params(self._h, None, ptr(p), NPARAM))

    def forward(self, F):
        F = np.ascontiguousarray(F, dtype=np.float32)
        y = np.zeros(len(F), np.float32)
        check(lib().gp_mlp_forward(self._h, ptr(F), len(F), ptr(y)))
        return y

    def gradient(self, F, t):
        F = np.ascontiguousarray(F, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_float()
        check(lib().gp_mlp_train(self._h, ptr(F), ptr(t), len(F), 0, 0, 0.0, 0, ptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, F, t, steps, batch=8192, lr=2e-3, seed=0):
        F = np.ascontiguousarray(F, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        check(lib().gp_mlp_train(self._h, ptr(F), ptr(t), len(F), steps, batch, lr, seed, None, C.byref(loss)))
        return loss.value

    def predict(self, mesh, sources):
        """dhat float64[V] and the device ms of features + inference."""
        src = _u32(sources)
        d = np.zeros(mesh.n_vertices())
        ms = C.c_float()
        check(lib().gp_mlp_predict(self._h, mesh.handle, ptr(src), len(src), ptr(d), C.byref(ms)))
        return d, ms.value

    def close(self):
        if getattr(self, "_h", None):
            lib().gp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def unpack(p):
    """FP32 parameter vector -> dict matching oracle/mlp.py."""
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (64, 32)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("w3", (64,)), ("b3", (1,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[k], np.float32).reshape(-1) for k in ("W1", "b1", "W2", "b2", "w3", "b3")])
