"""The reference driver's cylinder case end to end on the GPU against the CPU
oracle: O-grid metrics formed on the GPU from the mesh (transformed-periodic
wrap in theta, one-sided closures at the wall and the outer boundary), the
adiabatic no-slip wall on the cylinder, the cylindrical sponge outside,
uniform inflow.  Tolerances: metrics 1e-9 of the largest |J| for J and of the
largest |J_ij| for the J_ij (nested one-sided closures, as for the box meshes), sponge strength 1e-13, the
step 1e-9 relative L2 after 10 steps on identical metrics (momenta on the
scale of the momentum vector)."""
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import cylinder, hit
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

CONS = ("rho", "rhoU", "rhoV", "rhoW", "rhoE")
MET = ("J", "J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32", "J33")


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


@pytest.mark.parametrize("fd,filt,rk", [("PENTA10", "COMPACT10", "KENRK4"), ("PADE6", "COMPACT8", "RK4")])
def test_cylinder_case_matches_the_oracle(gpu, fd, filt, rk):
    from mpi_3dcompact_amd.solver import Solver
    Nx, Ny, Nz = 32, 56, 12
    x, y, z = cylinder.cylinder_mesh(Nx, Ny, Nz)
    rmin = 0.6
    sponge = {"KIND": "CYLINDRICAL", "STRENGTH": 1.0, "CYLAXIS_ORIENT": 2, "CYLAXIS_X": 0.0,
              "CYLAXIS_Y": 0.0, "CYLAXIS_Z": 0.0, "RMIN": rmin}
    case = dict(fd=fd, filt=filt, rk=rk, cfl=0.5, mu_ref=1e-3, faces=cylinder.FACES, dims=(Nx, Ny, Nz))
    _say(f"case: cylinder {Nx}x{Ny}x{Nz} {fd}/{filt}/{rk}, 10 steps")
    g = Solver(hit.options_text(Nx, sponge=sponge, disp=cylinder.DISP, **case), (Nz, Ny, Nx))
    g.set_mesh(x, y, z)
    ic = cylinder.initial_condition((Nz, Ny, Nx))
    g.set_initial(ic)
    # metrics from the mesh: GPU against the oracle's restatement
    fdc = [{"CD2": 0, "PADE6": 1, "PENTA10": 2}[fd]] * 3
    om = P.metrics(fdc, (0, 1, 0), x, y, z, cylinder.DISP)
    gm = [g.field(k) for k in MET]
    sj, sm = np.max(np.abs(om[0])), max(np.max(np.abs(m)) for m in om[1:])
    worst = max(np.max(np.abs(a - b)) / (sj if k == 0 else sm) for k, (a, b) in enumerate(zip(gm, om)))
    _say(f"  metrics: {worst:.2e} of the largest value (J apart from J_ij)")
    assert worst <= 1e-9
    # the metric identities the reference prints after forming the metrics
    # (src/AbstractSingleBlockMesh.cpp:1336-1393): plain first derivatives of
    # the J_ij, summed (xi + eta) + zeta, against the oracle's derivatives
    bc = (P.PERIODIC, P.DIRICHLET, P.PERIODIC)
    dd = (1.0 / Nx, 1.0 / (Ny - 1), 1.0 / Nz)
    got = g.metric_identity()
    for k in range(3):
        terms = [P.deriv_field(fdc[0], d, bc[d], gm[1 + 3 * d + k], dd[d]) for d in range(3)]
        ident = (terms[0] + terms[1]) + terms[2]
        scale = max(np.max(np.abs(t)) for t in terms)
        assert abs(got[k][0] - ident.min()) <= 1e-12 * scale and abs(got[k][1] - ident.max()) <= 1e-12 * scale, k
    # the cylindrical sponge strength from the same mesh
    fk = [P.FACE[k] for k in hit.face_kinds(cylinder.FACES)]
    sig = P.sponge_sigma(x, y, z, fk, kind=1, strength=1.0, cyl_orient=2, axis=(0.0, 0.0, 0.0), rmin=rmin)
    assert np.max(np.abs(g.field("sigma") - sig)) <= 1e-13 * max(1.0, np.max(sig))
    assert np.max(sig) > 0.1 and np.all(sig[:, :Ny // 2] == 0)
    # the step on identical metrics
    o = P.Solver(hit.oracle_config(Nx, **case), gm, ic)
    o.set_sponge(sig)
    for f in ("rho", "rhoU", "rhoE", "U", "T"):
        assert rel_l2(g.field(f), o.field(f)) <= 1e-13, ("initial", f)
    g.step(10)
    o.steps(10)
    # the flow is two-dimensional: rhoW is rounding noise, so the momenta are
    # compared on the scale of the whole momentum vector
    mom = np.sqrt(sum(np.linalg.norm(o.field(n)) ** 2 for n in ("rhoU", "rhoV", "rhoW")))
    errs = {n: (np.linalg.norm(g.field(n) - o.field(n)) / mom if n in ("rhoU", "rhoV", "rhoW")
                else rel_l2(g.field(n), o.field(n))) for n in CONS + ("p", "T")}
    assert np.max(np.abs(o.field("rhoW"))) <= 1e-10 * mom
    _say("  rel L2:", {k: f"{v:.2e}" for k, v in errs.items()})
    for n, e in errs.items():
        assert e <= 1e-9, (n, e)
    for f in ("U", "V", "W"):
        assert np.all(g.field(f)[:, 0, :] == 0), f   # no slip on the cylinder
    t, dt, ts = g.scalars()
    assert ts == o.timestep and abs(dt - o.dt) <= 1e-12 * o.dt
