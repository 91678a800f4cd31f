"""ctypes binding of libmc3d.so (the C ABI declared in include/mc3d.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C mpi_3dcompact_amd/csrc``).  There is no fallback: every entry point
runs the HIP implementation, and a missing library raises immediately.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmc3d.so")

MC3D_PERIODIC, MC3D_DIRICHLET = 0, 1
MC3D_XPENCIL, MC3D_YPENCIL, MC3D_ZPENCIL = 0, 1, 2
MC3D_CD2, MC3D_PADE6, MC3D_PENTA10 = 0, 1, 2
MC3D_COMPACT8, MC3D_COMPACT10, MC3D_COMMN5PADE, MC3D_COMMN5EXP = 0, 1, 2, 3

_VP = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)

# (name, restype, argtypes) for every exported entry point
SIGNATURES = [
    ("mc3d_last_error", ctypes.c_char_p, []),
    ("mc3d_device_count", ctypes.c_int, [_I]),
    ("mc3d_deriv_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _I,
                                         ctypes.c_double, ctypes.c_int, ctypes.POINTER(_VP)]),
    ("mc3d_deriv_apply", ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                                        ctypes.c_int, _VP]),
    ("mc3d_deriv_apply_tpb", ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(_VP), _VP]),
    ("mc3d_filter_create", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, _I, _I,
                                          ctypes.c_int, ctypes.POINTER(_VP)]),
    ("mc3d_filter_apply", ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                                         ctypes.c_int, _VP]),
    ("mc3d_lineop_destroy", None, [_VP]),
    ("mc3d_lineop_info", ctypes.c_int, [_VP, _I, _I, _I]),
    ("mc3d_lineop_kernel", ctypes.c_char_p, [_VP]),
    ("mc3d_dump_read_header", ctypes.c_int, [ctypes.c_char_p, _I, _D, _I]),
    ("mc3d_dump_read", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_D)]),
    ("mc3d_dump_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, _I,
                                       ctypes.POINTER(_D)]),
    ("mc3d_plan_tables", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int, _I, _D, _D,
                                        _D, _D, _D, _D, _I, _I]),
]

_LIB = None


class Mc3dError(RuntimeError):
    pass


def lib():
    """Load libmc3d.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise Mc3dError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                            "or `make -C mpi_3dcompact_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().mc3d_last_error().decode(errors="replace")
        raise Mc3dError(f"mc3d error {rc}: {msg}")
    return rc
