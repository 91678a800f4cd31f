"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker.  The product package never imports
this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

CD2, PADE6, PENTA10 = 0, 1, 2
COMPACT8, COMPACT10, COMMN5PADE, COMMN5EXP = 0, 1, 2, 3
PERIODIC, DIRICHLET = 0, 1

_D = ctypes.POINTER(ctypes.c_double)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_deriv_line.restype = ctypes.c_int
        L.orc_deriv_line.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     _D, _D, _D]
        L.orc_filter_line.restype = ctypes.c_int
        L.orc_filter_line.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      _D, _D]
        L.orc_neumann.restype = ctypes.c_double
        L.orc_neumann.argtypes = [ctypes.c_int, _D]
        L.orc_deriv_field.restype = ctypes.c_int
        L.orc_deriv_field.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D]
        L.orc_deriv_field_tpb.restype = ctypes.c_int
        L.orc_deriv_field_tpb.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, _D, _D,
                                          ctypes.POINTER(_D)]
        L.orc_filter_field.restype = ctypes.c_int
        L.orc_filter_field.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, _D, _D]
        L.orc_solve_tri.argtypes = [_D, _D, _D, _D, _D, _D, ctypes.c_int]
        L.orc_cyclic_tri.argtypes = [_D, _D, _D, ctypes.c_double, ctypes.c_double, _D,
                                     ctypes.c_int, _D]
        L.orc_solve_penta.argtypes = [_D, _D, _D, _D, _D, _D, _D, ctypes.c_int]
        L.orc_cyclic_penta.argtypes = [_D, _D, _D, _D, _D, _D, _D, _D, ctypes.c_int]
        L.orc_invert4.argtypes = [_D, _D]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_D)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def deriv_line(scheme, bc, f, dd, tpb=None):
    f = _c(f)
    out = np.empty_like(f)
    t = None if tpb is None else _c(tpb)
    rc = lib().orc_deriv_line(scheme, bc, f.size, dd, _p(f), _p(out), None if t is None else _p(t))
    if rc:
        raise ValueError(f"orc_deriv_line rc={rc}")
    return out


def filter_line(kind, bc, f, alpha_f=0.0):
    f = _c(f)
    out = np.empty_like(f)
    rc = lib().orc_filter_line(kind, bc, f.size, alpha_f, _p(f), _p(out))
    if rc:
        raise ValueError(f"orc_filter_line rc={rc}")
    return out


def neumann(scheme, f):
    f = _c(f)
    return lib().orc_neumann(scheme, _p(f))


def deriv_field(scheme, direction, bc, f, dd):
    """f has shape (Nz, Ny, Nx) (x fastest)."""
    f = _c(f)
    out = np.empty_like(f)
    Nz, Ny, Nx = f.shape
    rc = lib().orc_deriv_field(scheme, direction, bc, Nx, Ny, Nz, dd, _p(f), _p(out))
    if rc:
        raise ValueError(f"orc_deriv_field rc={rc}")
    return out


def deriv_field_tpb(scheme, direction, f, dd, wrap):
    """wrap: list of 6 arrays (Nm3,Nm2,Nm1,Np1,Np2,Np3), each one value per line."""
    f = _c(f)
    out = np.empty_like(f)
    Nz, Ny, Nx = f.shape
    ws = [_c(w).ravel() for w in wrap]
    arr = (_D * 6)(*[_p(w) for w in ws])
    rc = lib().orc_deriv_field_tpb(scheme, direction, Nx, Ny, Nz, dd, _p(f), _p(out), arr)
    if rc:
        raise ValueError(f"orc_deriv_field_tpb rc={rc}")
    return out


def filter_field(kind, alpha_f, direction, bc3, f):
    f = _c(f)
    out = np.empty_like(f)
    Nz, Ny, Nx = f.shape
    b = (ctypes.c_int * 3)(*bc3)
    rc = lib().orc_filter_field(kind, alpha_f, direction, b, Nx, Ny, Nz, _p(f), _p(out))
    if rc:
        raise ValueError(f"orc_filter_field rc={rc}")
    return out


def solve_tri(a, b, c, d):
    a, b, c, d = map(_c, (a, b, c, d))
    x = np.empty_like(d)
    w = np.empty_like(d)
    lib().orc_solve_tri(_p(a), _p(b), _p(c), _p(d), _p(x), _p(w), d.size)
    return x


def cyclic_tri(a, b, c, alpha, beta, r):
    a, b, c, r = map(_c, (a, b, c, r))
    x = np.empty_like(r)
    lib().orc_cyclic_tri(_p(a), _p(b), _p(c), alpha, beta, _p(r), r.size, _p(x))
    return x


def solve_penta(e, a, d, c, f, b):
    e, a, d, c, f, b = map(_c, (e, a, d, c, f, b))
    x = np.empty_like(b)
    lib().orc_solve_penta(_p(e), _p(a), _p(d), _p(c), _p(f), _p(b), _p(x), b.size)
    return x


def cyclic_penta(a, b, c, d, e, f, cp):
    a, b, c, d, e, f, cp = map(_c, (a, b, c, d, e, f, cp))
    x = np.empty_like(f)
    lib().orc_cyclic_penta(_p(a), _p(b), _p(c), _p(d), _p(e), _p(f), _p(cp), _p(x), f.size)
    return x


def invert4(m):
    m = _c(m).reshape(16)
    out = np.empty(16)
    lib().orc_invert4(_p(m), _p(out))
    return out.reshape(4, 4)
