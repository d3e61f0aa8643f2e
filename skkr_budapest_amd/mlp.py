"""ctypes binding of libskkr_mlp.so (include/skkr_mlp.h): the 3-layer
error-prediction MLP of the adaptive BZ mesh, trained on the GPU."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libskkr_mlp.so")
NHID = 64
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        V, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        L.skkr_mlp_create.argtypes = [ctypes.POINTER(V), I, I, I, ctypes.c_ulonglong]
        L.skkr_mlp_destroy.argtypes = [V]
        L.skkr_mlp_destroy.restype = None
        L.skkr_mlp_nparams.argtypes = [V]
        L.skkr_mlp_get_params.argtypes = [V, V]
        L.skkr_mlp_set_params.argtypes = [V, V]
        L.skkr_mlp_predict.argtypes = [V, I, V, V]
        L.skkr_mlp_loss_grad.argtypes = [V, I, V, V, V, V]
        L.skkr_mlp_train.argtypes = [V, I, V, V, I, F, V]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Mlp:
    def __init__(self, nin, seed=0, device=0):
        h = ctypes.c_void_p()
        rc = lib().skkr_mlp_create(ctypes.byref(h), nin, NHID, device, seed)
        if rc:
            raise RuntimeError(f"skkr_mlp_create failed ({rc})")
        self._h, self.nin = h, nin
        self.np = lib().skkr_mlp_nparams(h)
        self.shift = np.zeros(nin, np.float32)  # input standardisation
        self.scale = np.ones(nin, np.float32)

    def close(self):
        if getattr(self, "_h", None):
            lib().skkr_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.np, np.float32)
        lib().skkr_mlp_get_params(self._h, _p(p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        lib().skkr_mlp_set_params(self._h, _p(p))

    def _x(self, x):
        return np.ascontiguousarray((np.asarray(x, np.float32) - self.shift) / self.scale)

    def predict(self, x):
        xs = self._x(x)
        y = np.zeros(len(xs), np.float32)
        if len(xs) and lib().skkr_mlp_predict(self._h, len(xs), _p(xs), _p(y)):
            raise RuntimeError("skkr_mlp_predict failed")
        return y

    def loss_grad(self, x, t):
        xs = self._x(x)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.np, np.float32)
        if lib().skkr_mlp_loss_grad(self._h, len(xs), _p(xs), _p(t), _p(loss), _p(g)):
            raise RuntimeError("skkr_mlp_loss_grad failed")
        return float(loss[0]), g

    def fit(self, x, t, steps=2000, lr=3e-3):
        x = np.asarray(x, np.float32)
        self.shift = x.mean(0).astype(np.float32)
        self.scale = (x.std(0) + 1e-6).astype(np.float32)
        xs = self._x(x)
        t = np.ascontiguousarray(t, np.float32)
        hist = np.zeros(steps, np.float32)
        if lib().skkr_mlp_train(self._h, len(xs), _p(xs), _p(t), steps, lr, _p(hist)):
            raise RuntimeError("skkr_mlp_train failed")
        return hist
