"""SolutionDump files (the reference's restart / output format), host only.

Five doubles {timeStep, time, Nx, Ny, Nz}, then x, y, z, rho, rhoU, rhoV,
rhoW, rhoE as Nx*Ny*Nz doubles each, x fastest (arrays here are (Nz, Ny, Nx)).
Read and written by libmc3d.so (mc3d_dump_*).
"""
import ctypes

import numpy as np

from ._lib import check, lib

NAMES = ("x", "y", "z", "rho", "rhoU", "rhoV", "rhoW", "rhoE")
_D = ctypes.POINTER(ctypes.c_double)


def read_header(path):
    ts, t = ctypes.c_int(), ctypes.c_double()
    dims = (ctypes.c_int * 3)()
    check(lib().mc3d_dump_read_header(path.encode(), ctypes.byref(ts), ctypes.byref(t), dims))
    return ts.value, t.value, tuple(dims)


def read(path):
    """dict with timestep, time, dims and the eight arrays."""
    ts, t, dims = read_header(path)
    shape = (dims[2], dims[1], dims[0])
    arrs = [np.empty(shape) for _ in NAMES]
    ptrs = (_D * 8)(*[a.ctypes.data_as(_D) for a in arrs])
    check(lib().mc3d_dump_read(path.encode(), ptrs))
    out = dict(zip(NAMES, arrs))
    out.update(timestep=ts, time=t, dims=dims)
    return out


def write(path, timestep, time, fields):
    """fields: dict name -> (Nz, Ny, Nx) array; missing names are written as zeros."""
    shape = next(iter(fields.values())).shape
    dims = (ctypes.c_int * 3)(shape[2], shape[1], shape[0])
    keep = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in fields.items()}
    ptrs = (_D * 8)(*[keep[n].ctypes.data_as(_D) if n in keep else None for n in NAMES])
    check(lib().mc3d_dump_write(path.encode(), int(timestep), float(time), dims, ptrs))
