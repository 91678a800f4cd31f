"""Python handle over the ensemble C ABI (mc3d_ensemble_*): many periodic
boxes of uneven sizes advanced together, members of equal size sharing every
launch, one HIP graph per step after the first."""
import ctypes
import json

import numpy as np

from ._lib import check, lib

_VP = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)

_SIGS = [
    ("mc3d_ensemble_create", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _I, _D, ctypes.POINTER(_VP)]),
    ("mc3d_ensemble_set_metrics", ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(_D)]),
    ("mc3d_ensemble_set_initial", ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _D, _D, _D]),
    ("mc3d_ensemble_tune", ctypes.c_int, [_VP, ctypes.c_char_p]),
    ("mc3d_ensemble_step", ctypes.c_int, [_VP, ctypes.c_int]),
    ("mc3d_ensemble_use_graphs", ctypes.c_int, [_VP, ctypes.c_int]),
    ("mc3d_ensemble_get_field", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_char_p, _D]),
    ("mc3d_ensemble_scalars", ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _I]),
    ("mc3d_ensemble_stream", ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
    ("mc3d_ensemble_sync", ctypes.c_int, [_VP]),
    ("mc3d_ensemble_info", ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_int, _I]),
    ("mc3d_ensemble_destroy", None, [_VP]),
]


def _bind():
    L = lib()
    if not getattr(L, "_ensemble_bound", False):
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._ensemble_bound = True
    return L


def _p(a):
    return a.ctypes.data_as(_D)


class Ensemble:
    """dims: list of (nx, ny, nz) per member; mu_ref: list per member (None: the options' MU_REF).
    Fields are (nz, ny, nx) arrays of one member."""

    def __init__(self, options_text, dims, mu_ref=None):
        L = _bind()
        self.dims = [tuple(int(v) for v in d) for d in dims]
        d = np.ascontiguousarray(np.array(self.dims, dtype=np.int32).reshape(-1))
        mu = None if mu_ref is None else np.ascontiguousarray(np.asarray(mu_ref, dtype=np.float64))
        if mu is not None and mu.size != len(self.dims):
            raise ValueError("one MU_REF per member")
        h = _VP()
        check(L.mc3d_ensemble_create(options_text.encode(), len(self.dims), d.ctypes.data_as(_I),
                                     None if mu is None else _p(mu), ctypes.byref(h)))
        self._h = h
        self._L = L

    def __len__(self):
        return len(self.dims)

    def set_metrics(self, m, metrics):
        ms = [np.ascontiguousarray(x, dtype=np.float64) for x in metrics]
        arr = (_D * 10)(*[_p(x) for x in ms])
        check(self._L.mc3d_ensemble_set_metrics(self._h, m, arr))

    def set_initial(self, m, ic):
        a = [np.ascontiguousarray(ic[k], dtype=np.float64) for k in ("rho", "U", "V", "W", "p")]
        check(self._L.mc3d_ensemble_set_initial(self._h, m, *[_p(x) for x in a]))

    def tune(self, cache_path=None):
        """Measure and select the line-operator launch shapes (cache_path: text cache, read and updated)."""
        check(self._L.mc3d_ensemble_tune(self._h, None if cache_path is None else str(cache_path).encode()))

    def step(self, n=1):
        check(self._L.mc3d_ensemble_step(self._h, n))

    def use_graphs(self, on=True):
        check(self._L.mc3d_ensemble_use_graphs(self._h, 1 if on else 0))

    def field(self, m, name):
        nx, ny, nz = self.dims[m]
        out = np.empty((nz, ny, nx))
        check(self._L.mc3d_ensemble_get_field(self._h, m, name.encode(), _p(out)))
        return out

    def scalars(self, m):
        t, dt, ts = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(self._L.mc3d_ensemble_scalars(self._h, m, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(ts)))
        return t.value, dt.value, ts.value

    @property
    def stream(self):
        s = _VP()
        check(self._L.mc3d_ensemble_stream(self._h, ctypes.byref(s)))
        return s.value

    def sync(self):
        check(self._L.mc3d_ensemble_sync(self._h))

    def info(self):
        need = ctypes.c_int(0)
        buf = ctypes.create_string_buffer(1 << 16)
        rc = self._L.mc3d_ensemble_info(self._h, buf, len(buf), ctypes.byref(need))
        if rc != 0 and need.value > len(buf):
            buf = ctypes.create_string_buffer(need.value)
            rc = self._L.mc3d_ensemble_info(self._h, buf, len(buf), ctypes.byref(need))
        check(rc)
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.mc3d_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
