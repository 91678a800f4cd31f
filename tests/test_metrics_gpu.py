"""GPU metric terms (mc3d_mesh_metrics and the solver's set_mesh) against the
CPU oracle of AbstractSingleBlockMesh::solveForJacobians on the same meshes.

Tolerance: max |gpu - oracle| <= tol * scale, scale = max |J| for J and
the largest |J_ij| over all nine for each J_ij (some are zero).  The
metrics nest three derivatives (x -> V = x_xi y -> J_ij / J -> x J_i), so
rounding differences of the line solves (1e-16 relative) grow by about the
operator norm per level: ~1 / d(xi) for periodic lines (tol 1e-11) and
several times that at the one-sided PENTA10 closures (tol 1e-9, the bound the
oracle's own affine-exactness check needs there).  An index, sign, wrap-slot
or translation mistake is O(1).
"""
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

PER, DIR = P.PERIODIC, P.DIRICHLET
TOL = 1e-11


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def grid(N, bc):
    axes = [np.arange(N[d]) / N[d] if bc[d] == PER else np.arange(N[d]) / (N[d] - 1) for d in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return x, y, z


def affine(A, N, bc):
    xi = grid(N, bc)
    X = [A[i][0] * xi[0] + A[i][1] * xi[1] + A[i][2] * xi[2] for i in range(3)]
    trans = [[A[e][d] if bc[d] == PER else 0.0 for e in range(3)] for d in range(3)]
    return X, trans


def wavy(N, bc, amp, t):
    xi, eta, zeta = grid(N, bc)
    return [xi + amp * np.sin(t * eta) * np.sin(t * zeta),
            eta + amp * np.sin(t * zeta) * np.sin(t * xi),
            zeta + amp * np.sin(t * xi) * np.sin(t * eta)], [[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]]


A_GEN = np.array([[2.0, 0.3, 0.0], [-0.4, 1.5, 0.0], [0.25, 0.1, 1.2]])
A_TYPO = np.array([[1.0, 0.0, 0.0], [0.0, 1.3, 0.4], [0.3, 0.0, 0.9]])


def max_rel_errs(got, ref):
    sj = max(np.max(np.abs(ref[0])), 1e-300)
    sm = max(max(np.max(np.abs(r)) for r in ref[1:]), 1e-300)
    return [np.max(np.abs(g - r)) / (sj if k == 0 else sm) for k, (g, r) in enumerate(zip(got, ref))]


def tol_for(bc):
    return TOL if all(b == PER for b in bc) else 1e-9


def compare(torch, fd, bc, X, trans, label):
    from mpi_3dcompact_amd.mesh import mesh_metrics
    N = X[0].shape[::-1]
    _say(f"  metrics case {label} fd={fd} bc={bc} N={N}")
    ref = P.metrics(fd, bc, *X, trans)
    dev = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in X]
    got = [o.cpu().numpy() for o in mesh_metrics(fd, bc, *dev, trans)]
    errs = max_rel_errs(got, ref)
    _say("   max rel err per field:", " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= tol_for(bc), errs
    return got, ref


@pytest.mark.parametrize("fd", [P.PADE6, P.PENTA10])
@pytest.mark.parametrize("bc", [(PER, PER, PER), (PER, DIR, DIR), (DIR, DIR, PER), (DIR, DIR, DIR)])
def test_affine_mesh(torch, fd, bc):
    X, trans = affine(A_GEN, (16, 18, 20), bc)
    compare(torch, (fd,) * 3, bc, X, trans, "affine")


@pytest.mark.parametrize("fd", [P.PADE6, P.PENTA10])
def test_smooth_dirichlet_mesh(torch, fd):
    X, _ = wavy((24, 20, 22), (DIR,) * 3, 0.1, np.pi)
    compare(torch, (fd,) * 3, (DIR,) * 3, X, [[0] * 3] * 3, "wavy-dirichlet")


@pytest.mark.parametrize("fd", [P.PADE6, P.PENTA10])
def test_smooth_periodic_mesh(torch, fd):
    """non-affine periodic map: the transformed-periodic path with wrap values
    of products (its boundary-layer behaviour is the reference's own)."""
    X, trans = wavy((32, 32, 32), (PER,) * 3, 0.05, 2 * np.pi)
    compare(torch, (fd,) * 3, (PER,) * 3, X, trans, "wavy-periodic")


def test_seven_point_typo_slots(torch):
    """PENTA10 with a map whose y_zeta z product reaches the Nm3 slot typo."""
    X, trans = affine(A_TYPO, (16, 16, 16), (PER,) * 3)
    got, _ = compare(torch, (P.PENTA10,) * 3, (PER,) * 3, X, trans, "typo")
    Ainv = np.linalg.inv(A_TYPO)
    assert np.max(np.abs(got[4] - Ainv[1, 0])) > 1e-6  # the typo acts on the GPU too


def test_mixed_schemes(torch):
    X, trans = affine(A_GEN, (20, 16, 18), (PER, DIR, PER))
    compare(torch, (P.PENTA10, P.CD2, P.PADE6), (PER, DIR, PER), X, trans, "mixed")


def test_cd2_periodic_is_rejected(torch):
    from mpi_3dcompact_amd._lib import Mc3dError
    from mpi_3dcompact_amd.mesh import mesh_metrics
    X, trans = affine(A_GEN, (12, 12, 12), (PER,) * 3)
    dev = [torch.from_numpy(v).cuda() for v in X]
    with pytest.raises(Mc3dError):
        mesh_metrics((P.CD2,) * 3, (PER,) * 3, *dev, trans)


def test_solver_set_mesh_matches_oracle_run(torch):
    """A sheared periodic box x = A xi through the solver's set_mesh (metrics
    formed on the GPU from PERIODICDISP = columns of A), then 10 steps, against
    the oracle solver given the oracle's metrics of the same mesh."""
    from mpi_3dcompact_amd.solver import Solver
    N = 32
    A = 2 * np.pi * np.array([[1.0, 0.15, 0.0], [0.0, 1.0, 0.1], [0.05, 0.0, 1.0]])
    X, trans = affine(A, (N, N, N), (PER,) * 3)
    case = dict(fd="PENTA10", filt="COMPACT8", alpha_f=0.45, rk="KENRK4", les="VREMAN")
    disp = [[trans[d][e] for e in range(3)] for d in range(3)]
    g = Solver(hit.options_text(N, disp=disp, **case), (N, N, N))
    g.set_mesh(*X)
    ref_m = P.metrics((P.PENTA10,) * 3, (PER,) * 3, *X, trans)
    errs = max_rel_errs([g.field(n) for n in ("J", "J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32",
                                              "J33")], ref_m)
    _say("   set_mesh max rel err:", " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= TOL, errs
    # the metric identities (src/AbstractSingleBlockMesh.cpp:1336-1393) of these
    # metrics against the oracle's plain derivatives of the same fields: the
    # transformed-periodic metrics of the sheared box vary near the wrap (the
    # reference's own behaviour), so the identities are not zero here
    gm = [g.field(n) for n in ("J", "J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32", "J33")]
    got = g.metric_identity()
    for k in range(3):
        terms = [P.deriv_field(P.PENTA10, d, PER, gm[1 + 3 * d + k], 1.0 / N) for d in range(3)]
        ident = (terms[0] + terms[1]) + terms[2]
        scale = max(np.max(np.abs(t)) for t in terms)
        assert abs(got[k][0] - ident.min()) <= 1e-12 * scale and abs(got[k][1] - ident.max()) <= 1e-12 * scale, k
    # constant metrics: the identities vanish to rounding
    h = Solver(hit.options_text(N, **case), (N, N, N))
    h.set_metrics(hit.identity_metrics(N))
    assert max(max(abs(lo), abs(hi)) for lo, hi in h.metric_identity()) <= 1e-13
    ic = hit.hit_initial_condition(N, k0=3.0, Mt=0.2, seed=4)
    g.set_initial(ic)
    g.step(10)
    o = P.Solver(hit.oracle_config(N, **case), ref_m, ic)
    o.steps(10)
    for name in ("rho", "rhoU", "rhoV", "rhoW", "rhoE"):
        a, b = g.field(name), o.field(name)
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        _say(f"   {name} rel L2 {e:.2e}")
        assert e <= 1e-9, (name, e)
