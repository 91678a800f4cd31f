"""Python handles over the derivative / filter operators of libmc3d.so.

Fields are torch tensors on the GPU, float64, contiguous, passed to the C ABI
as raw device pointers.  In the default layout (MC3D_XPENCIL) a field has
shape (Nz, Ny, Nx), x fastest; in the reference's Y pencil (MC3D_YPENCIL)
shape (Nx, Nz, Ny), y fastest; in its Z pencil (MC3D_ZPENCIL) shape
(Ny, Nx, Nz), z fastest.
"""
import ctypes

import numpy as np

from ._lib import (MC3D_PERIODIC, MC3D_XPENCIL, check, lib)

_VP = ctypes.c_void_p


def _ptr(t):
    if not t.is_cuda:
        raise ValueError("field must be a GPU tensor")
    if not t.is_contiguous():
        raise ValueError("field must be contiguous")
    import torch
    if t.dtype != torch.float64:
        raise ValueError("field must be float64")
    return t.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


class LineOp:
    """A derivative (``kind='deriv'``) or filter (``kind='filter'``) along one direction."""

    def __init__(self, handle, cls, dims, direction, layout=MC3D_XPENCIL):
        self._h = handle
        self.cls = cls
        self.dims = tuple(dims)
        self.direction = direction
        self.layout = layout

    @classmethod
    def deriv(cls, scheme, direction, bc, dims, dd, layout=MC3D_XPENCIL):
        h = _VP()
        d = (ctypes.c_int * 3)(*dims)
        check(lib().mc3d_deriv_create(scheme, direction, bc, d, dd, layout, ctypes.byref(h)))
        return cls(h, "deriv", dims, direction, layout)

    @classmethod
    def filter(cls, kind, alpha_f, direction, bc3, dims, layout=MC3D_XPENCIL):
        h = _VP()
        d = (ctypes.c_int * 3)(*dims)
        b = (ctypes.c_int * 3)(*bc3)
        check(lib().mc3d_filter_create(kind, alpha_f, direction, b, d, layout, ctypes.byref(h)))
        return cls(h, "filter", dims, direction, layout)

    def shape(self):
        """Tensor shape of a field in this operator's layout (slowest axis first)."""
        r = self.layout
        fast = [self.dims[(m + r) % 3] for m in range(3)]
        return (fast[2], fast[1], fast[0])

    def info(self):
        p, m, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().mc3d_lineop_info(self._h, ctypes.byref(p), ctypes.byref(m), ctypes.byref(l)))
        return p.value, m.value, l.value

    def kernel(self):
        """Kernel instantiation a plain application launches (profiler name)."""
        return lib().mc3d_lineop_kernel(self._h).decode()

    def apply(self, ins, outs, stream=None):
        ins = list(ins) if isinstance(ins, (list, tuple)) else [ins]
        outs = list(outs) if isinstance(outs, (list, tuple)) else [outs]
        n = len(ins)
        if len(outs) != n:
            raise ValueError("ins and outs differ in length")
        shape = self.shape()
        for t in ins + outs:
            if tuple(t.shape) != shape:
                raise ValueError(f"field shape {tuple(t.shape)} != {shape}")
        pi = (_VP * n)(*[_ptr(t) for t in ins])
        po = (_VP * n)(*[_ptr(t) for t in outs])
        fn = lib().mc3d_deriv_apply if self.cls == "deriv" else lib().mc3d_filter_apply
        check(fn(self._h, pi, po, n, _stream_ptr(stream)))

    def apply_tpb(self, inp, out, wraps, stream=None):
        w = (_VP * 6)(*[(_ptr(x) if x is not None else None) for x in wraps])
        check(lib().mc3d_deriv_apply_tpb(self._h, _ptr(inp), _ptr(out), w, _stream_ptr(stream)))

    def close(self):
        if self._h:
            lib().mc3d_lineop_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_tables(cls, kind, bc, n, dd=1.0, alpha_f=0.0, mm=0):
    """Host-built partition tables of an operator (diagnostic; no GPU needed)."""
    sizes = (ctypes.c_int * 11)()
    check(lib().mc3d_plan_tables(cls, kind, bc, n, dd, alpha_f, mm, sizes, None, None, None, None,
                                 None, None, None, None))
    N, W, H, P, NI, nbl, nbr, int_off, MM, dense_ok, gb = list(sizes)
    W2 = max(2 * W, 1)
    nrow = max(nbl + nbr, 1)
    lu = np.zeros(P * MM * 5)
    sp = np.zeros(P * MM * W2)
    G = np.zeros(P * W2 * max(NI, 1))
    rows = np.zeros(nrow * 19)
    drows = np.zeros(nrow * 17)
    interior = np.zeros(11)
    cstart = np.zeros(P, dtype=np.int32)
    clen = np.zeros(P, dtype=np.int32)
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    check(lib().mc3d_plan_tables(cls, kind, bc, n, dd, alpha_f, mm, sizes,
                                 lu.ctypes.data_as(D), sp.ctypes.data_as(D), G.ctypes.data_as(D),
                                 rows.ctypes.data_as(D), drows.ctypes.data_as(D),
                                 interior.ctypes.data_as(D),
                                 cstart.ctypes.data_as(I), clen.ctypes.data_as(I)))
    return dict(N=N, W=W, H=H, P=P, NI=NI, nbl=nbl, nbr=nbr, int_off=int_off, MM=MM,
                dense_ok=bool(dense_ok), gb=gb,
                lu=lu.reshape(P, MM, 5), sp=sp.reshape(P, MM, W2), G=G.reshape(P, W2, max(NI, 1)),
                rows=rows.reshape(-1, 19), drows=drows.reshape(-1, 17), interior=interior,
                cstart=cstart, clen=clen, periodic=(bc == MC3D_PERIODIC))
