"""Python front end of the warm-start MLP library (include/usg_mlp.h, in libusg.so).

No CPU fallback: the network runs through the gfx950 BF16 MFMA kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import usg

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)

EXPORTS = ["usg_mlp_sizes", "usg_mlp_create", "usg_mlp_destroy", "usg_mlp_set_params", "usg_mlp_get_params",
           "usg_mlp_predict", "usg_mlp_gradients", "usg_mlp_set_data", "usg_mlp_train_epoch",
           "usg_mlp_mfma_check", "usg_mlp_last_error"]


class MlpError(RuntimeError):
    pass


class _Desc(C.Structure):
    _fields_ = [("nin", C.c_int), ("width", C.c_int), ("nlayer", C.c_int), ("nout", C.c_int)]


_bound = False


def lib():
    global _bound
    L = usg.lib()
    if not _bound:
        vp = C.c_void_p
        L.usg_mlp_sizes.argtypes = [C.POINTER(_Desc), C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.usg_mlp_create.argtypes = [C.POINTER(_Desc), C.POINTER(vp)]
        L.usg_mlp_destroy.argtypes = [vp]
        L.usg_mlp_set_params.argtypes = [vp, _fp, _fp]
        L.usg_mlp_get_params.argtypes = [vp, _fp, _fp]
        L.usg_mlp_predict.argtypes = [vp, _fp, C.c_long, _fp]
        L.usg_mlp_gradients.argtypes = [vp, _fp, _fp, C.c_int, _fp, _fp, _fp]
        L.usg_mlp_set_data.argtypes = [vp, _fp, _fp, C.c_long]
        L.usg_mlp_train_epoch.argtypes = [vp, _ip, C.c_long, C.c_int, C.c_float, _fp]
        L.usg_mlp_mfma_check.argtypes = [_fp, _fp, _fp]
        L.usg_mlp_last_error.restype = C.c_char_p
        _bound = True
    return L


def _check(rc):
    if rc != 0:
        raise MlpError(f"usg_mlp error {rc}: {lib().usg_mlp_last_error().decode()}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def sizes(nin, width, nlayer, nout):
    d = _Desc(nin, width, nlayer, nout)
    nw, nb = C.c_long(), C.c_long()
    _check(lib().usg_mlp_sizes(C.byref(d), C.byref(nw), C.byref(nb)))
    return nw.value, nb.value


def mfma_check(a, b):
    a, b = _f32(a), _f32(b)
    d = np.zeros((16, 16), np.float32)
    _check(lib().usg_mlp_mfma_check(a.ctypes.data_as(_fp), b.ctypes.data_as(_fp), d.ctypes.data_as(_fp)))
    return d


class MLP:
    def __init__(self, nin, width=256, nlayer=8, nout=4):
        self.nin, self.width, self.nlayer, self.nout = nin, width, nlayer, nout
        self.nw, self.nb = sizes(nin, width, nlayer, nout)
        h = C.c_void_p()
        _check(lib().usg_mlp_create(C.byref(_Desc(nin, width, nlayer, nout)), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().usg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, w, b):
        w, b = _f32(w), _f32(b)
        assert w.size == self.nw and b.size == self.nb
        _check(lib().usg_mlp_set_params(self._h, w.ctypes.data_as(_fp), b.ctypes.data_as(_fp)))

    def get_params(self):
        w, b = np.zeros(self.nw, np.float32), np.zeros(self.nb, np.float32)
        _check(lib().usg_mlp_get_params(self._h, w.ctypes.data_as(_fp), b.ctypes.data_as(_fp)))
        return w, b

    def predict(self, x):
        x = _f32(x).reshape(-1, self.nin)
        y = np.zeros((x.shape[0], self.nout), np.float32)
        _check(lib().usg_mlp_predict(self._h, x.ctypes.data_as(_fp), x.shape[0], y.ctypes.data_as(_fp)))
        return y

    def gradients(self, x, y):
        x, y = _f32(x).reshape(-1, self.nin), _f32(y).reshape(-1, self.nout)
        loss = C.c_float()
        dw, db = np.zeros(self.nw, np.float32), np.zeros(self.nb, np.float32)
        _check(lib().usg_mlp_gradients(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp), x.shape[0],
                                       C.byref(loss), dw.ctypes.data_as(_fp), db.ctypes.data_as(_fp)))
        return loss.value, dw, db

    def set_data(self, x, y):
        self._x, self._y = _f32(x).reshape(-1, self.nin), _f32(y).reshape(-1, self.nout)
        _check(lib().usg_mlp_set_data(self._h, self._x.ctypes.data_as(_fp), self._y.ctypes.data_as(_fp),
                                      self._x.shape[0]))
        self.nset = self._x.shape[0]

    def train_epoch(self, perm, batch, lr):
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        loss = C.c_float()
        _check(lib().usg_mlp_train_epoch(self._h, perm.ctypes.data_as(_ip), perm.size, batch, lr, C.byref(loss)))
        return loss.value
