"""Learned SGS closure: the network handle over the mc3d_closure_* C ABI.

The closure replaces the dynamic procedure of the compressible DSM
(src/DSMSGS.hpp:124-444) behind the AbstractSGS interface: per point, seven
invariant features of the resolved velocity gradient and the filter width go
through a 7-64-64-64-2 GELU network whose outputs are the DSM's coefficient C
and C / Pr_t (mu_sgs = rho C Delta^2 |S|, k_sgs = cp (C / Pr_t) rho Delta^2 |S|).
Training data and network inputs are torch float32 tensors on the GPU:
X (n, 7) standardised features, T (n, 2) standardised targets.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .ops import _stream_ptr

_VP = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)
_F7 = ctypes.c_float * 7
_F2 = ctypes.c_float * 2

N_IN, N_OUT = 7, 2
N_PARAMS = 8962

_SIGS = [
    ("mc3d_closure_create", ctypes.c_int, [ctypes.c_ulonglong, ctypes.POINTER(_VP)]),
    ("mc3d_closure_destroy", None, [_VP]),
    ("mc3d_closure_set_params", ctypes.c_int, [_VP, _F]),
    ("mc3d_closure_get_params", ctypes.c_int, [_VP, _F]),
    ("mc3d_closure_set_scaling", ctypes.c_int, [_VP, _F, _F, _F, _F, _F, _F]),
    ("mc3d_closure_get_scaling", ctypes.c_int, [_VP, _F, _F, _F, _F, _F, _F]),
    ("mc3d_closure_gradient", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_long, _VP,
                                             ctypes.POINTER(ctypes.c_double), _VP]),
    ("mc3d_closure_adam_step", ctypes.c_int, [_VP, _VP, ctypes.c_float, _VP]),
    ("mc3d_closure_train", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_ulonglong,
                                          ctypes.POINTER(ctypes.c_double), _VP]),
    ("mc3d_closure_infer", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_long, _VP]),
    ("mc3d_closure_save", ctypes.c_int, [_VP, ctypes.c_char_p]),
    ("mc3d_closure_load", ctypes.c_int, [_VP, ctypes.c_char_p]),
]


def _bind():
    L = lib()
    if not getattr(L, "_closure_bound", False):
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._closure_bound = True
    return L


def _fptr(t, cols):
    import torch
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous() or t.dim() != 2 or t.shape[1] != cols:
        raise ValueError(f"expected a contiguous float32 GPU tensor of shape (n, {cols})")
    return t.data_ptr()


class Closure:
    """The closure network with its Adam state."""

    def __init__(self, seed=0):
        self._L = _bind()
        h = _VP()
        check(self._L.mc3d_closure_create(seed, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.mc3d_closure_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def params(self):
        p = np.empty(N_PARAMS, np.float32)
        check(self._L.mc3d_closure_get_params(self._h, p.ctypes.data_as(_F)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters")
        check(self._L.mc3d_closure_set_params(self._h, p.ctypes.data_as(_F)))

    def set_scaling(self, in_mean, in_std, out_mean, out_std, lo, hi):
        a = [(_F7 if len(v) == N_IN else _F2)(*[float(x) for x in v])
             for v in (in_mean, in_std, out_mean, out_std, lo, hi)]
        check(self._L.mc3d_closure_set_scaling(self._h, *a))

    def scaling(self):
        a = [_F7(), _F7(), _F2(), _F2(), _F7(), _F7()]
        check(self._L.mc3d_closure_get_scaling(self._h, *a))
        keys = ("in_mean", "in_std", "out_mean", "out_std", "lo", "hi")
        return {k: np.array(v[:], np.float32) for k, v in zip(keys, a)}

    def gradient(self, X, T, stream=None):
        """Full-batch gradient (torch float32 GPU tensor of N_PARAMS) and loss."""
        import torch
        g = torch.empty(N_PARAMS, dtype=torch.float32, device=X.device)
        loss = ctypes.c_double()
        check(self._L.mc3d_closure_gradient(self._h, _fptr(X, N_IN), _fptr(T, N_OUT), X.shape[0], g.data_ptr(),
                                            ctypes.byref(loss), _stream_ptr(stream)))
        return g, loss.value

    def adam_step(self, grad, lr=1e-3, stream=None):
        check(self._L.mc3d_closure_adam_step(self._h, grad.data_ptr(), lr, _stream_ptr(stream)))

    def train(self, X, T, epochs=20, batch=8192, lr=1e-3, seed=0, stream=None):
        """Shuffled minibatch Adam; returns the mean loss of every epoch."""
        losses = (ctypes.c_double * epochs)()
        check(self._L.mc3d_closure_train(self._h, _fptr(X, N_IN), _fptr(T, N_OUT), X.shape[0], epochs, batch, lr,
                                         seed, losses, _stream_ptr(stream)))
        return list(losses)

    def infer(self, X, stream=None):
        import torch
        Y = torch.empty((X.shape[0], N_OUT), dtype=torch.float32, device=X.device)
        check(self._L.mc3d_closure_infer(self._h, _fptr(X, N_IN), Y.data_ptr(), X.shape[0], _stream_ptr(stream)))
        return Y

    def save(self, path):
        check(self._L.mc3d_closure_save(self._h, str(path).encode()))

    def load(self, path):
        check(self._L.mc3d_closure_load(self._h, str(path).encode()))
