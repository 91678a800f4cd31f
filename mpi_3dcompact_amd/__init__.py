"""MI355X-native (gfx950) per-timestep hot path of a compact-difference
curvilinear compressible flow solver (derivatives, filters, flux/RHS, RK stage
updates, SGS models), behind the C ABI of include/mc3d.h.

Python here is a thin ctypes view used by tests and benchmarks; the product is
libmc3d.so (HIP kernels + C ABI + C++ host classes).
"""
from ._lib import Mc3dError, check, lib  # noqa: F401
from . import ops  # noqa: F401
