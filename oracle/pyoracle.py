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
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_max_threads.argtypes = []
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def max_threads():
    """Threads the oracle's OpenMP loops run with (omp_get_max_threads)."""
    return lib().orc_max_threads()


def set_threads(n):
    """Threads for the oracle's later OpenMP loops (omp_set_num_threads)."""
    lib().orc_set_threads(int(n))


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


def metrics(fd, bc, x, y, z, trans):
    """J, J11..J33 (list of 10 arrays shaped like x) from mesh coordinates."""
    L = lib()
    if not getattr(L, "_metrics_bound", False):
        L.orc_metrics.restype = ctypes.c_int
        L.orc_metrics.argtypes = [ctypes.POINTER(ctypes.c_int)] * 3 + [ctypes.POINTER(ctypes.c_double)] + \
            [_D] * 3 + [ctypes.POINTER(_D)]
        L._metrics_bound = True
    shape = x.shape
    N = (ctypes.c_int * 3)(shape[2], shape[1], shape[0])
    fdc = (ctypes.c_int * 3)(*fd)
    bcc = (ctypes.c_int * 3)(*bc)
    tr = (ctypes.c_double * 9)(*[float(t) for row in trans for t in row])
    xs = [_c(np.ascontiguousarray(a, dtype=np.float64)) for a in (x, y, z)]
    outs = [np.empty(shape) for _ in range(10)]
    arr = (_D * 10)(*[o.ctypes.data_as(_D) for o in outs])
    rc = L.orc_metrics(fdc, bcc, N, tr, *[a.ctypes.data_as(_D) for a in xs], arr)
    if rc:
        raise RuntimeError(f"orc_metrics failed ({rc})")
    return outs


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


# ---------------------------------------------------------------------------
# solver step oracle (orc_solver.h)
# ---------------------------------------------------------------------------
TVDRK3, RK4, KENRK4, LSLDDRK4 = 0, 1, 2, 3
NOMODEL, VREMAN, DSM = 0, 1, 2
AVG_NONE, AVG_XI1, AVG_XI2, AVG_XI3, AVG_LOCAL, AVG_GLOBAL = 0, 1, 2, 3, 4, 5
CONST_DT, CONST_CFL = 0, 1

FIELD_IDS = {name: i for i, name in enumerate(
    ["rho", "rhoU", "rhoV", "rhoW", "rhoE", "U", "V", "W", "p", "T", "mu", "sos",
     "mu_sgs", "taukk", "k_sgs", "C", "CI", "Prt",
     "rhok", "rhoUk", "rhoVk", "rhoWk", "rhoEk", "rhok2", "rhoUk2", "rhoVk2", "rhoWk2", "rhoEk2",
     "UAVG", "VAVG", "WAVG", "URMS", "VRMS", "WRMS", "UVREY", "UWREY", "VWREY",
     "RHOAVG", "RHORMS", "PAVG", "PRMS", "TAVG", "TRMS", "rho2"])}


class OrcConfig(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
                ("bc", ctypes.c_int * 3), ("fd", ctypes.c_int * 3), ("filter", ctypes.c_int),
                ("alphaF", ctypes.c_double), ("rk", ctypes.c_int), ("les", ctypes.c_int),
                ("les_avg", ctypes.c_int), ("dsm_test_filter", ctypes.c_int),
                ("dsm_use_taukk", ctypes.c_int), ("dsm_coeff_range", ctypes.c_int),
                ("mu_ref", ctypes.c_double), ("ts_type", ctypes.c_int), ("cfl", ctypes.c_double),
                ("dt", ctypes.c_double), ("filter_step", ctypes.c_int),
                ("substep_filtering", ctypes.c_int), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("timestep0", ctypes.c_int),
                ("time0", ctypes.c_double), ("face", ctypes.c_int * 6),
                ("velocity_stats", ctypes.c_int), ("thermo_stats", ctypes.c_int),
                ("stats_avg", ctypes.c_int), ("stats_interval", ctypes.c_int)]


# face kinds (src/Options.hpp BCKind order)
FACE = {"PERIODIC": 0, "INTERNALLY_PERIODIC": 1, "SPONGE": 2, "ADIABATIC_WALL": 3, "CONST_T_WALL": 4,
        "MOVING_ADIABATIC_WALL": 5, "INLET": 6}


def sponge_sigma(x, y, z, faces, kind=0, strength=1.0, perc=(0,) * 6, cyl_orient=2, axis=(0, 0, 0), rmin=0.0):
    """Sponge strength field of src/SpongeBC.cpp from mesh coordinates ((Nz, Ny, Nx) arrays)."""
    L = lib()
    L.orc_sponge_sigma.restype = ctypes.c_int
    Nz, Ny, Nx = x.shape
    xs = [_c(a) for a in (x, y, z)]
    out = np.empty_like(xs[0])
    I6, D6, D3 = ctypes.c_int * 6, ctypes.c_double * 6, ctypes.c_double * 3
    rc = L.orc_sponge_sigma(ctypes.c_int(Nx), ctypes.c_int(Ny), ctypes.c_int(Nz), *[_p(a) for a in xs],
                            I6(*faces), ctypes.c_int(kind), ctypes.c_double(strength), D6(*perc),
                            ctypes.c_int(cyl_orient), D3(*axis), ctypes.c_double(rmin), _p(out))
    if rc:
        raise ValueError("unsupported sponge configuration")
    return out


def _solver_lib():
    L = lib()
    if not hasattr(L, "_solver_bound"):
        L.orc_solver_create.restype = ctypes.c_void_p
        L.orc_solver_create.argtypes = [ctypes.POINTER(OrcConfig), ctypes.POINTER(_D), _D, _D, _D,
                                        _D, _D]
        L.orc_solver_steps.restype = ctypes.c_int
        L.orc_solver_steps.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_solver_field.restype = _D
        L.orc_solver_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        for nm in ("orc_solver_time", "orc_solver_dt"):
            getattr(L, nm).restype = ctypes.c_double
            getattr(L, nm).argtypes = [ctypes.c_void_p]
        L.orc_solver_stats_weight.restype = ctypes.c_double
        L.orc_solver_stats_weight.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_solver_timestep.restype = ctypes.c_int
        L.orc_solver_timestep.argtypes = [ctypes.c_void_p]
        L.orc_solver_destroy.restype = None
        L.orc_solver_destroy.argtypes = [ctypes.c_void_p]
        L.orc_solver_rhs.restype = ctypes.c_int
        L.orc_solver_rhs.argtypes = [ctypes.c_void_p, ctypes.POINTER(_D)]
        L.orc_rk_amplification.restype = ctypes.c_double
        L.orc_rk_amplification.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L._solver_bound = True
    return L


class Solver:
    """CPU oracle of the solver.  cfg: dict of OrcConfig fields; metrics: list of
    10 arrays (J, J11..J33); ic: dict rho, U, V, W, p.  Arrays shaped (Nz,Ny,Nx)."""

    def __init__(self, cfg, metrics, ic):
        L = _solver_lib()
        c = OrcConfig()
        for k, v in cfg.items():
            if k in ("bc", "fd", "face"):
                getattr(c, k)[:] = list(v)
            else:
                setattr(c, k, v)
        self.shape = (cfg["Nz"], cfg["Ny"], cfg["Nx"])
        self._keep = [_c(m) for m in metrics] + [_c(ic[k]) for k in ("rho", "U", "V", "W", "p")]
        mp = (_D * 10)(*[_p(m) for m in self._keep[:10]])
        r, u, v, w, p = self._keep[10:]
        self._h = L.orc_solver_create(ctypes.byref(c), mp, _p(r), _p(u), _p(v), _p(w), _p(p))
        self._L = L

    def set_sponge(self, sigma, avg=None):
        """sigma field; avg: dict rho, rhoU, rhoV, rhoW, rhoE (default: the current state)."""
        self._keep.append(_c(sigma))
        self._L.orc_solver_set_sponge.restype = ctypes.c_int
        arr = None
        if avg is not None:
            a = [_c(avg[k]) for k in ("rho", "rhoU", "rhoV", "rhoW", "rhoE")]
            self._keep += a
            arr = (_D * 5)(*[_p(v) for v in a])
        self._L.orc_solver_set_sponge(ctypes.c_void_p(self._h), _p(self._keep[-6 if avg is not None else -1]), arr)

    def steps(self, n):
        self._L.orc_solver_steps(self._h, n)

    def field(self, name):
        ptr = self._L.orc_solver_field(self._h, FIELD_IDS[name])
        if not ptr:
            raise KeyError(f"{name} is not kept by this configuration")
        n = self.shape[0] * self.shape[1] * self.shape[2]
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.shape).copy()

    def rhs(self):
        n = self.shape[0] * self.shape[1] * self.shape[2]
        outs = [np.empty(n) for _ in range(5)]
        arr = (_D * 5)(*[_p(o) for o in outs])
        self._L.orc_solver_rhs(self._h, arr)
        return [o.reshape(self.shape) for o in outs]

    def stats_weight(self, which):
        """accumulated statistics weight: 0 velocity, 1 thermo"""
        return self._L.orc_solver_stats_weight(self._h, which)

    @property
    def time(self):
        return self._L.orc_solver_time(self._h)

    @property
    def dt(self):
        return self._L.orc_solver_dt(self._h)

    @property
    def timestep(self):
        return self._L.orc_solver_timestep(self._h)

    def __del__(self):
        try:
            self._L.orc_solver_destroy(self._h)
        except Exception:
            pass


def rk_amplification(rk, z, dual_step=1):
    return _solver_lib().orc_rk_amplification(rk, dual_step, z)
