"""Python handles over the SGS-model and CFL entry points of libmc3d.so.

``Sgs`` mirrors the reference's AbstractSGS (src/AbstractSGS.hpp:35-41) over
caller-owned GPU fields: torch float64 tensors, contiguous, in the handle's
layout (MC3D_XPENCIL: shape (Nz, Ny, Nx); MC3D_YPENCIL, the reference's SGS
layout: shape (Nx, Nz, Ny); MC3D_ZPENCIL: (Ny, Nx, Nz)).  ``cfl_dt`` is
calcDtFromCFL (src/CurvilinearCSolver_Core.cpp:189-237).
"""
import ctypes

from ._lib import MC3D_XPENCIL, check, lib

_VP = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_L = ctypes.POINTER(ctypes.c_long)

_SIGS = [
    ("mc3d_cfl_dt", ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_long, ctypes.c_int, _D, ctypes.c_double, _VP,
                                   _VP]),
    ("mc3d_cfl_dt_noalloc", ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_long, ctypes.c_int, _D, ctypes.c_double,
                                           _VP, _VP]),
    ("mc3d_sgs_create", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _VP, _VP, ctypes.POINTER(_VP)]),
    ("mc3d_sgs_destroy", None, [_VP]),
    ("mc3d_sgs_vreman", ctypes.c_int, [_VP, ctypes.POINTER(_VP), _VP, _VP]),
    ("mc3d_sgs_dsm_mu_taukk", ctypes.c_int, [_VP, ctypes.POINTER(_VP), _VP, _VP, _VP, _VP, _VP, _VP]),
    ("mc3d_sgs_dsm_ksgs", ctypes.c_int, [_VP, ctypes.POINTER(_VP), _VP, _VP, _VP, _VP, _VP, _VP]),
    ("mc3d_sgs_coefficients", ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ("mc3d_sgs_set_closure", ctypes.c_int, [_VP, _VP]),
    ("mc3d_sgs_learned", ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), _VP, _VP, _VP, _VP, _VP,
                                        _VP, _VP]),
    ("mc3d_sgs_closure_stats", ctypes.c_int, [_VP, _L, _L, _L, _L]),
]


def _bind():
    L = lib()
    if not getattr(L, "_sgs_bound", False):
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._sgs_bound = True
    return L


def _ptr(t):
    import torch
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous() or t.dtype != torch.float64:
        raise ValueError("fields must be contiguous float64 GPU tensors")
    return t.data_ptr()


def _ptrs(ts):
    return (_VP * len(ts))(*[_ptr(t) for t in ts])


def _stream(stream):
    import torch
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def cfl_dt(U, V, W, sos, J11, J22, J33, spacing, cfl, members=1, stream=None, out=None, noalloc=False):
    """(dt, maxspeed) per box as a (members, 2) float64 GPU tensor (written
    into out when given).  noalloc: mc3d_cfl_dt_noalloc, which allocates
    nothing and returns without waiting for the stream."""
    import torch
    n = U.numel()
    if n % members:
        raise ValueError("points not divisible by members")
    if out is None:
        out = torch.empty((members, 2), dtype=torch.float64, device=U.device)
    elif out.numel() < 2 * members:
        raise ValueError("out holds fewer than 2 * members doubles")
    sp = (ctypes.c_double * 3)(*spacing)
    fn = _bind().mc3d_cfl_dt_noalloc if noalloc else _bind().mc3d_cfl_dt
    check(fn(_ptrs([U, V, W, sos, J11, J22, J33]), n // members, members, sp, cfl, _ptr(out), _stream(stream)))
    return out


class Sgs:
    """An SGS model (VREMAN, DSM or LEARNED per the options text) over
    caller-owned fields; J is the Jacobian field (kept by reference)."""

    def __init__(self, options_text, J, layout=MC3D_XPENCIL, stream=None):
        self._h = _VP()
        self._J = J  # the handle keeps the pointer
        check(_bind().mc3d_sgs_create(options_text.encode(), layout, _ptr(J), _stream(stream),
                                      ctypes.byref(self._h)))

    def vreman(self, gradU, rho, mu_sgs):
        check(_bind().mc3d_sgs_vreman(self._h, _ptrs(gradU), _ptr(rho), _ptr(mu_sgs)))

    def dsm_mu_taukk(self, gradU, rho, rhoU, rhoV, rhoW, mu_sgs, taukk):
        check(_bind().mc3d_sgs_dsm_mu_taukk(self._h, _ptrs(gradU), _ptr(rho), _ptr(rhoU), _ptr(rhoV),
                                            _ptr(rhoW), _ptr(mu_sgs), _ptr(taukk)))

    def dsm_ksgs(self, gradT, rho, rhoU, rhoV, rhoW, T, k_sgs):
        check(_bind().mc3d_sgs_dsm_ksgs(self._h, _ptrs(gradT), _ptr(rho), _ptr(rhoU), _ptr(rhoV), _ptr(rhoW),
                                        _ptr(T), _ptr(k_sgs)))

    def coefficients(self, C=None, CI=None, Prt=None):
        """Copy C, CI, Prt of the last DSM call (DSMCOEFFRANGE) into the given fields."""
        check(_bind().mc3d_sgs_coefficients(self._h, _ptr(C), _ptr(CI), _ptr(Prt)))

    def set_closure(self, closure):
        check(_bind().mc3d_sgs_set_closure(self._h, closure._h))

    def learned(self, gradU, rho, mu_sgs, k_sgs, gradT=None, rhoU=None, rhoV=None, rhoW=None, T=None):
        gT = _ptrs(gradT) if gradT is not None else None
        check(_bind().mc3d_sgs_learned(self._h, _ptrs(gradU), gT, _ptr(rho), _ptr(rhoU), _ptr(rhoV), _ptr(rhoW),
                                       _ptr(T), _ptr(mu_sgs), _ptr(k_sgs)))

    def closure_stats(self):
        v = [ctypes.c_long() for _ in range(4)]
        check(_bind().mc3d_sgs_closure_stats(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def close(self):
        if self._h:
            _bind().mc3d_sgs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
