"""Metric terms of a single-block curvilinear mesh on the GPU
(AbstractSingleBlockMesh::solveForJacobians, src/AbstractSingleBlockMesh.cpp:23-1309).

Coordinates are torch float64 GPU tensors shaped (Nz, Ny, Nx); the ten results
{J, J11 .. J33} come back the same way.  Computational coordinates run over
[0, 1) in a periodic direction and [0, 1] otherwise; a periodic direction's
coordinates jump by ``trans[d]`` across its boundary (PERIODICDISP rows,
:2019-2085).
"""
import ctypes

from ._lib import check, lib
from .ops import _ptr, _stream_ptr

_VP = ctypes.c_void_p
_I3 = ctypes.c_int * 3
_D9 = ctypes.c_double * 9

METRIC_NAMES = ("J", "J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32", "J33")


def _bind():
    L = lib()
    if not getattr(L, "_mesh_bound", False):
        fn = L.mc3d_mesh_metrics
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.POINTER(ctypes.c_int)] * 3 + [ctypes.POINTER(ctypes.c_double), _VP, _VP, _VP,
                                                            ctypes.POINTER(_VP), _VP]
        L._mesh_bound = True
    return L


def mesh_metrics(fd, bc, x, y, z, trans=None, stream=None):
    """fd, bc: per-direction scheme (MC3D_CD2/PADE6/PENTA10) and boundary
    (MC3D_PERIODIC/DIRICHLET); trans: 3x3 translations or None.  Returns the
    list of 10 metric tensors."""
    import torch
    L = _bind()
    if not (x.shape == y.shape == z.shape) or x.dim() != 3:
        raise ValueError("x, y, z must be 3-D fields of one shape")
    Nz, Ny, Nx = x.shape
    out = [torch.empty_like(x) for _ in range(10)]
    tr = None if trans is None else _D9(*[float(t) for row in trans for t in row])
    ptrs = (_VP * 10)(*[_ptr(o) for o in out])
    check(L.mc3d_mesh_metrics(_I3(*fd), _I3(*bc), _I3(Nx, Ny, Nz), tr, _ptr(x), _ptr(y), _ptr(z), ptrs,
                              _stream_ptr(stream)))
    return out
