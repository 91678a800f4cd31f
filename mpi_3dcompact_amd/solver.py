"""Python handle over the whole-solver C ABI (mc3d_solver_*)."""
import ctypes

import numpy as np

from ._lib import check, lib

_VP = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)

_SIGS = [
    ("mc3d_options_check", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]),
    ("mc3d_solver_create", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_VP)]),
    ("mc3d_solver_set_metrics", ctypes.c_int, [_VP, ctypes.POINTER(_D)]),
    ("mc3d_solver_set_mesh", ctypes.c_int, [_VP, _D, _D, _D]),
    ("mc3d_solver_set_initial", ctypes.c_int, [_VP, _D, _D, _D, _D, _D]),
    ("mc3d_solver_set_sponge", ctypes.c_int, [_VP, _D, ctypes.POINTER(_D)]),
    ("mc3d_solver_step", ctypes.c_int, [_VP, ctypes.c_int]),
    ("mc3d_solver_load_restart", ctypes.c_int, [_VP]),
    ("mc3d_solver_write_dump", ctypes.c_int, [_VP, ctypes.c_char_p, _D, _D, _D]),
    ("mc3d_solver_set_output", ctypes.c_int, [_VP, ctypes.c_char_p, _D, _D, _D]),
    ("mc3d_solver_run", ctypes.c_int, [_VP]),
    ("mc3d_solver_get_field", ctypes.c_int, [_VP, ctypes.c_char_p, _D]),
    ("mc3d_solver_field_device", ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.POINTER(_VP)]),
    ("mc3d_solver_scalars", ctypes.c_int, [_VP, _D, _D, ctypes.POINTER(ctypes.c_int)]),
    ("mc3d_solver_stream", ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
    ("mc3d_solver_sync", ctypes.c_int, [_VP]),
    ("mc3d_solver_destroy", None, [_VP]),
    ("mc3d_solver_timing", ctypes.c_int, [_VP, ctypes.c_int]),
    ("mc3d_solver_set_closure", ctypes.c_int, [_VP, _VP]),
    ("mc3d_solver_sample_closure", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_long, ctypes.c_long, ctypes.c_long]),
    ("mc3d_solver_eval_sgs", ctypes.c_int, [_VP]),
    ("mc3d_solver_hit_stats", ctypes.c_int, [_VP, _D]),
    ("mc3d_solver_set_turbdata", ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_int]),
    ("mc3d_solver_stats_weights", ctypes.c_int, [_VP, _D, _D]),
    ("mc3d_solver_add_image", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                             ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    ("mc3d_solver_metric_identity", ctypes.c_int, [_VP, _D]),
    ("mc3d_solver_slice", ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D]),
    ("mc3d_image_write", ctypes.c_int, [ctypes.c_char_p, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int]),
    ("mc3d_solver_closure_stats", ctypes.c_int, [_VP] + [ctypes.POINTER(ctypes.c_long)] * 4),
    ("mc3d_solver_timing_report", ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int)]),
]


def _bind():
    L = lib()
    if not getattr(L, "_solver_bound", False):
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._solver_bound = True
    return L


def options_check(text):
    """Parse and validate an options text without a device; returns the parsed
    values as a {"SECTION.KEY": "value"} dict (raises Mc3dError with the
    reference's complaint, MC3D_ECONFIG, on a bad deck)."""
    L = _bind()
    need = ctypes.c_int(0)
    check(L.mc3d_options_check(text.encode(), None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    check(L.mc3d_options_check(text.encode(), buf, need.value, None))
    out = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition(" = ")
        out[k] = v
    return out


def _p(a):
    return a.ctypes.data_as(_D)


COLORMAPS = {"GREYSCALE": 0, "RAINBOW": 1, "BWR": 2}


def image_write(path, plane, width, height, vrange=None, colormap="RAINBOW"):
    """Draw an (n2, n1) plane as the solver's image outputs do and write the PNG
    (host only)."""
    a = np.ascontiguousarray(plane, dtype=np.float64)
    cm = COLORMAPS[colormap] if isinstance(colormap, str) else int(colormap)
    lo, hi = vrange if vrange is not None else (0.0, 1.0)
    check(_bind().mc3d_image_write(str(path).encode(), _p(a), a.shape[1], a.shape[0], width, height,
                                   vrange is not None, lo, hi, cm))


class Solver:
    """GPU solver configured by options-file text; fields are (Nz, Ny, Nx) arrays."""

    def __init__(self, options_text, shape):
        L = _bind()
        h = _VP()
        check(L.mc3d_solver_create(options_text.encode(), ctypes.byref(h)))
        self._h = h
        self._L = L
        self.shape = tuple(shape)

    def set_metrics(self, metrics):
        ms = [np.ascontiguousarray(m, dtype=np.float64) for m in metrics]
        arr = (_D * 10)(*[_p(m) for m in ms])
        check(self._L.mc3d_solver_set_metrics(self._h, arr))

    def set_mesh(self, x, y, z):
        """Metrics from mesh coordinates (AbstractSingleBlockMesh::solveForJacobians)."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
        check(self._L.mc3d_solver_set_mesh(self._h, *[_p(v) for v in a]))

    def set_sponge(self, sigma, avg=None):
        """Sponge strength field; avg: dict rho, rhoU, rhoV, rhoW, rhoE of the sponge
        averages (default: the state at set_initial)."""
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        arr = None
        keep = [sg]
        if avg is not None:
            a = [np.ascontiguousarray(avg[k], dtype=np.float64) for k in ("rho", "rhoU", "rhoV", "rhoW", "rhoE")]
            keep += a
            arr = (_D * 5)(*[_p(v) for v in a])
        check(self._L.mc3d_solver_set_sponge(self._h, _p(sg), arr))

    def set_closure(self, closure):
        """Use a trained mpi_3dcompact_amd.closure.Closure (LESMODEL LEARNED)."""
        check(self._L.mc3d_solver_set_closure(self._h, closure.handle))

    def sample_closure(self, X, T, count, offset=0, stride=1):
        """Training pairs of the current DSM state into float32 GPU tensors X (>= count, 7), T (>= count, 2)."""
        if X.shape[0] < count or T.shape[0] < count:
            raise ValueError("sample buffers too small")
        check(self._L.mc3d_solver_sample_closure(self._h, X.data_ptr(), T.data_ptr(), count, offset, stride))

    def eval_sgs(self):
        """Derivatives and SGS fields of the current state (not advanced)."""
        check(self._L.mc3d_solver_eval_sgs(self._h))

    HIT_STATS = ("time", "tau", "re_lambda", "mach_t", "dissipation", "uprime", "kolmogorov", "rhoprime",
                 "dilprime", "vortprime", "rho_uu")

    def hit_stats(self):
        """Decaying-HIT statistics (the reference's turbdata.out columns) as a dict."""
        out = (ctypes.c_double * 11)()
        check(self._L.mc3d_solver_hit_stats(self._h, out))
        return dict(zip(self.HIT_STATS, out[:]))

    def set_turbdata(self, path, every=25):
        """Append the hit_stats line to `path` (turbdata.out) every `every` steps; None: off."""
        check(self._L.mc3d_solver_set_turbdata(self._h, None if path is None else str(path).encode(), every))

    def add_image(self, field, var_name, plane=2, fraction=0.5, width=512, height=512, interval=0,
                  vrange=None, colormap="RAINBOW"):
        """PngWriter image output: <var_name>.<step>.png of the index plane nearest
        `fraction` along `plane`, every `interval` steps (0: the first step only),
        colour range vrange = (min, max) or the plane's extremes."""
        cm = COLORMAPS[colormap] if isinstance(colormap, str) else int(colormap)
        lo, hi = vrange if vrange is not None else (0.0, 1.0)
        check(self._L.mc3d_solver_add_image(self._h, interval, width, height, field.encode(), var_name.encode(),
                                            plane, fraction, vrange is not None, lo, hi, cm))

    def metric_identity(self):
        """(min, max) of the three metric identities of the current metrics."""
        out = (ctypes.c_double * 6)()
        check(self._L.mc3d_solver_metric_identity(self._h, out))
        return [(out[2 * k], out[2 * k + 1]) for k in range(3)]

    def slice(self, field, direction, index):
        """Index plane of a field normal to `direction`, shaped (n2, n1) with the
        image axes: direction 0 -> (k, j), 1 -> (i, k), 2 -> (j, i)."""
        nz, ny, nx = self.shape
        n1, n2 = {0: (ny, nz), 1: (nz, nx), 2: (nx, ny)}[direction]
        out = np.empty((n2, n1))
        check(self._L.mc3d_solver_slice(self._h, field.encode(), direction, index, _p(out)))
        return out

    def stats_weights(self):
        """Accumulated weights (time covered) of the velocity and thermo statistics."""
        v, t = ctypes.c_double(), ctypes.c_double()
        check(self._L.mc3d_solver_stats_weights(self._h, ctypes.byref(v), ctypes.byref(t)))
        return v.value, t.value

    def closure_stats(self):
        v = [ctypes.c_long() for _ in range(4)]
        check(self._L.mc3d_solver_closure_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("evaluations", "outside_points", "fallback_stages", "fallback_points"), [x.value for x in v]))

    def set_initial(self, ic):
        a = [np.ascontiguousarray(ic[k], dtype=np.float64) for k in ("rho", "U", "V", "W", "p")]
        check(self._L.mc3d_solver_set_initial(self._h, *[_p(x) for x in a]))

    def load_restart(self):
        """FROMRESTART: initial state from the options file's RESTART.FILENAME."""
        check(self._L.mc3d_solver_load_restart(self._h))

    def _mesh(self, mesh):
        if mesh is None:
            return [None, None, None], []
        arrs = [np.ascontiguousarray(m, dtype=np.float64) for m in mesh]
        return [_p(a) for a in arrs], arrs

    def write_dump(self, path=None, mesh=None):
        """SolutionDump of the current state; mesh = (x, y, z) host arrays or None."""
        ptrs, keep = self._mesh(mesh)
        check(self._L.mc3d_solver_write_dump(self._h, path.encode() if path else None, *ptrs))

    def set_output(self, directory=None, mesh=None):
        """Dump every DUMPSTEP steps and at the end of run() into directory."""
        ptrs, keep = self._mesh(mesh)
        check(self._L.mc3d_solver_set_output(self._h, directory.encode() if directory else None, *ptrs))

    def step(self, n=1):
        check(self._L.mc3d_solver_step(self._h, n))

    def run(self):
        check(self._L.mc3d_solver_run(self._h))

    def field(self, name):
        out = np.empty(self.shape)
        check(self._L.mc3d_solver_get_field(self._h, name.encode(), _p(out)))
        return out

    def device_ptr(self, name):
        p = _VP()
        check(self._L.mc3d_solver_field_device(self._h, name.encode(), ctypes.byref(p)))
        return p.value

    def scalars(self):
        t, dt, ts = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(self._L.mc3d_solver_scalars(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(ts)))
        return t.value, dt.value, ts.value

    def stream(self):
        p = _VP()
        check(self._L.mc3d_solver_stream(self._h, ctypes.byref(p)))
        return p.value

    def sync(self):
        check(self._L.mc3d_solver_sync(self._h))

    def timing(self, enable=True):
        check(self._L.mc3d_solver_timing(self._h, 1 if enable else 0))

    def timing_report(self):
        import json
        need = ctypes.c_int(0)
        self._L.mc3d_solver_timing_report(self._h, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(max(need.value, 2))
        check(self._L.mc3d_solver_timing_report(self._h, buf, len(buf), ctypes.byref(need)))
        return json.loads(buf.value.decode())

    def close(self):
        if self._h:
            self._L.mc3d_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
