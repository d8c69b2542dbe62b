"""The far-field closure network on the MI355X: 32 -> 128 -> 128 -> 128 -> 16
SiLU MLP, BF16 weights and activations on MFMA with FP32 accumulation, with
hand-written forward, backward and Adam kernels (fv_mlp_*).

Parameters live on the device as FP32 masters (layout W1[128][32], b1[128],
W2[128][128], b2, W3[128][128], b3, W4[16][128], b4[16]) plus BF16 copies used
by the MFMA kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

N_IN, N_OUT = 32, 16
SHAPES = [("W1", (128, N_IN)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)),
          ("W3", (128, 128)), ("b3", (128,)), ("W4", (N_OUT, 128)), ("b4", (N_OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def init_params(n_in=N_IN, n_out=N_OUT, seed=0, zero_output=False) -> np.ndarray:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases; padded
    inputs / outputs beyond n_in / n_out get zero weights.  zero_output starts
    the output layer at zero (a residual network that initially predicts 0)."""
    rng = np.random.default_rng(seed)
    out = []
    fan = {"W1": n_in, "W2": 128, "W3": 128, "W4": 128}
    for name, shp in SHAPES:
        if name.startswith("W"):
            a = 1.0 / np.sqrt(fan[name])
            w = rng.uniform(-a, a, shp)
            if name == "W1":
                w[:, n_in:] = 0.0
            if name == "W4":
                w[n_out:, :] = 0.0
                if zero_output:
                    w[:] = 0.0
            out.append(w.ravel())
        else:
            out.append(np.zeros(int(np.prod(shp))))
    return np.concatenate(out).astype(np.float32)


def closure_params(seed=0) -> np.ndarray:
    """Initial parameters of the far-field closure: 31 inputs (the features of
    fv_closure.hip), 3 outputs (the velocity residual), zero output layer."""
    return init_params(n_in=31, n_out=3, seed=seed, zero_output=True)


class ClosureMLP:
    def __init__(self, params=None, context=None, seed=0):
        self._ctx = context or _lib.default_context()
        h = C.c_void_p()
        self._ctx.check(self._ctx.lib.fv_mlp_create(self._ctx.handle, C.byref(h)))
        self._h = h
        assert self._ctx.lib.fv_mlp_param_count() == NPARAMS
        self.set_params(init_params(seed=seed) if params is None else params)

    def __del__(self):
        try:
            if self._h and self._h.value:
                self._ctx.lib.fv_mlp_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (NPARAMS,):
            raise ValueError(f"expected {NPARAMS} parameters, got {p.shape}")
        self._ctx.check(self._ctx.lib.fv_mlp_set_params(self._h, p.ctypes.data_as(_lib._fp)))

    def get_params(self) -> np.ndarray:
        p = np.zeros(NPARAMS, dtype=np.float32)
        self._ctx.check(self._ctx.lib.fv_mlp_get_params(self._h, p.ctypes.data_as(_lib._fp)))
        return p

    @staticmethod
    def _pad(X, width=N_IN):
        X = np.asarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], width), dtype=np.float32)
        out[:, :X.shape[1]] = X
        return out

    def forward(self, X) -> np.ndarray:
        Xp = self._pad(X)
        Y = np.zeros((Xp.shape[0], N_OUT), dtype=np.float32)
        if Xp.shape[0]:
            self._ctx.check(self._ctx.lib.fv_mlp_forward_host(
                self._h, Xp.ctypes.data_as(_lib._fp), Xp.shape[0], Y.ctypes.data_as(_lib._fp)))
        return Y

    def gradients(self, X, T, wout) -> np.ndarray:
        Xp, Tp = self._pad(X), self._pad(T, N_OUT)
        w = self._pad(np.asarray(wout, dtype=np.float32)[None, :], N_OUT)[0]
        g = np.zeros(NPARAMS, dtype=np.float32)
        self._ctx.check(self._ctx.lib.fv_mlp_gradients_host(
            self._h, Xp.ctypes.data_as(_lib._fp), Tp.ctypes.data_as(_lib._fp),
            w.ctypes.data_as(_lib._fp), Xp.shape[0], g.ctypes.data_as(_lib._fp)))
        return g

    def adam(self, grad, lr):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        self._ctx.check(self._ctx.lib.fv_mlp_adam_host(self._h, g.ctypes.data_as(_lib._fp),
                                                       float(lr)))
