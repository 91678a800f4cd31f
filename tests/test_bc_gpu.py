"""Wall and sponge boundaries of the GPU step against the CPU oracle
(src/CurvilinearCSolver_BC.cpp:3-842, src/SpongeBC.cpp) on identical seeded
inputs.  Tolerance: relative L2 <= 1e-9 on every conserved field after the
stated steps (FP64), as for the periodic step; sponge strength fields to 1e-14.
"""
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

CONS = ("rho", "rhoU", "rhoV", "rhoW", "rhoE")


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def box_metrics(N, faces):
    """metrics of x = 2 pi xi with xi in [0, 1) periodic or [0, 1] otherwise: J_ii = 1 / (2 pi)
    whatever the faces, the grid spacing in xi following them."""
    return hit.identity_metrics(N)


def run_pair(N, steps, faces, sponge=None, seed=0, k0=3.0, Mt=0.2, **case):
    from mpi_3dcompact_amd.solver import Solver
    ic = hit.hit_initial_condition(N, k0=k0, Mt=Mt, seed=seed)
    met = box_metrics(N, faces)
    g = Solver(hit.options_text(N, faces=faces, sponge=sponge, **case), (N, N, N))
    o = P.Solver(hit.oracle_config(N, faces=faces, **case), met, ic)
    if sponge is not None:
        x, y, z = hit.box_mesh(N, faces)
        g.set_mesh(x, y, z)
        fk = [P.FACE[k] for k in hit.face_kinds(faces)]
        perc = [float(sponge.get(f"RECT{f}PERC", 0.0)) for f in hit.FACE_NAMES]
        sig = P.sponge_sigma(x, y, z, fk, kind=0, strength=float(sponge.get("STRENGTH", 1.0)), perc=perc)
        g.set_initial(ic)
        assert np.max(np.abs(g.field("sigma") - sig)) <= 1e-14 * max(1.0, np.max(np.abs(sig)))
        o.set_sponge(sig)
        # the mesh-formed metrics equal the box metrics to the rounding of the compact
        # derivatives of linear coordinates (~1e-11 relative); the oracle uses the latter
        for k, name in enumerate(("J", "J11", "J22", "J33")):
            ref = met[(0, 1, 5, 9)[k]]
            assert np.max(np.abs(g.field(name) - ref)) <= 1e-10 * np.max(np.abs(ref))
        g.set_metrics(met)
    else:
        g.set_metrics(met)
        g.set_initial(ic)
    for f in ("rho", "rhoU", "rhoE", "T", "p"):
        e = rel_l2(g.field(f), o.field(f))
        assert e <= 1e-13, ("initial", f, e)
    g.step(steps)
    o.steps(steps)
    return g, o


def check(g, o, tol=1e-9):
    errs = {n: rel_l2(g.field(n), o.field(n)) for n in CONS + ("T", "p")}
    _say("  rel L2:", {k: f"{v:.2e}" for k, v in errs.items()})
    for n, e in errs.items():
        assert e <= tol, (n, e)
    t, dt, ts = g.scalars()
    assert ts == o.timestep and abs(dt - o.dt) <= 1e-12 * o.dt


CASES = [
    dict(faces={"Y0": "ADIABATIC_WALL", "Y1": "ADIABATIC_WALL"}, fd="PADE6", filt="COMPACT10", rk="TVDRK3",
         mu_ref=2e-3, steps=8),
    dict(faces={"X0": "ADIABATIC_WALL", "X1": "ADIABATIC_WALL", "Z0": "MOVING_ADIABATIC_WALL",
                "Z1": "CONST_T_WALL"}, fd="PENTA10", filt="COMPACT8", rk="KENRK4", les="VREMAN",
         les_avg="LOCAL", mu_ref=1e-3, steps=5),
    dict(faces={"X0": "CONST_T_WALL", "X1": "CONST_T_WALL", "Y0": "ADIABATIC_WALL", "Y1": "ADIABATIC_WALL",
                "Z0": "ADIABATIC_WALL", "Z1": "MOVING_ADIABATIC_WALL"}, fd="CD2", filt="COMPACT8", rk="RK4",
         mu_ref=5e-3, cfl=0.3, steps=5),
    dict(faces={"Y0": "ADIABATIC_WALL", "Y1": "ADIABATIC_WALL"}, fd="PADE6", filt="COMPACT10", rk="KENRK4",
         les="DSM", les_avg="GLOBAL", mu_ref=2e-3, steps=3),
]


@pytest.mark.parametrize("case", CASES, ids=["walls-y", "walls-xz-vreman", "walls-all-cd2", "walls-y-dsm"])
def test_walls(gpu, case):
    case = dict(case)
    steps = case.pop("steps")
    _say(f"case: {case['faces']} {steps} steps")
    g, o = run_pair(32, steps, **case)
    check(g, o)
    for f in ("U", "V", "W"):
        v = g.field(f)
        for face, sl in (("Y0", (slice(None), 0)), ("Y1", (slice(None), -1))):
            if case["faces"].get(face, "PERIODIC") != "PERIODIC":
                assert np.all(v[sl] == 0), (f, face)


def test_sponge_and_walls(gpu):
    sponge = {"KIND": "RECTILINEAR", "STRENGTH": 2.0, "RECTX0PERC": 0.2, "RECTX1PERC": 0.15}
    faces = {"X0": "SPONGE", "X1": "SPONGE", "Y0": "ADIABATIC_WALL", "Y1": "ADIABATIC_WALL"}
    _say("case: sponge in x, walls in y")
    g, o = run_pair(32, 6, faces, sponge=sponge, fd="PADE6", filt="COMPACT10", rk="TVDRK3", mu_ref=2e-3)
    check(g, o)


def test_sponge_needs_its_strength(gpu):
    from mpi_3dcompact_amd._lib import Mc3dError
    from mpi_3dcompact_amd.solver import Solver
    N = 16
    faces = {"X0": "SPONGE", "X1": "SPONGE"}
    g = Solver(hit.options_text(N, faces=faces, sponge={"KIND": "RECTILINEAR", "RECTX0PERC": 0.2,
                                                        "RECTX1PERC": 0.2}), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(hit.hit_initial_condition(N, k0=2.0, Mt=0.1, seed=1))
    with pytest.raises(Mc3dError):
        g.step(1)
    g.set_sponge(np.zeros((N, N, N)))
    g.step(1)
