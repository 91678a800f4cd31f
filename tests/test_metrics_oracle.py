"""Metric terms from mesh coordinates (oracle restatement of
AbstractSingleBlockMesh::solveForJacobians, src/AbstractSingleBlockMesh.cpp:23-1439).

Pinned by what the mathematics fixes:
- an affine map x = A xi (any invertible A, periodic with translation A[:, d]
  across direction d, or Dirichlet) has constant metrics J_ij = d xi_i / d x_j
  = (A^-1)_ij and J = 1/det A, which compact derivatives of linear functions
  reproduce to rounding; a non-symmetric A catches transposed indices and
  sign mistakes;
- a smooth periodic non-affine map against its analytic inverse Jacobian, to
  the scheme's accuracy;
- the two 7-point call-site typos of the reference (T8) break the affine
  exactness exactly where they act, and only there.
"""
import numpy as np
import pytest

from oracle import pyoracle as P

PER, DIR = P.PERIODIC, P.DIRICHLET


def grid(N, bc):
    """computational coordinates xi_d in [0, 1) (periodic) or [0, 1] (Dirichlet)."""
    axes = []
    for d in range(3):
        n = N[d]
        axes.append(np.arange(n) / n if bc[d] == PER else np.arange(n) / (n - 1))
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return x, y, z  # xi, eta, zeta as (Nz, Ny, Nx) arrays


def affine(A, N, bc):
    xi, eta, zeta = grid(N, bc)
    X = [A[i][0] * xi + A[i][1] * eta + A[i][2] * zeta for i in range(3)]
    trans = [[A[e][d] if bc[d] == PER else 0.0 for e in range(3)] for d in range(3)]
    return X, trans


A_GEN = np.array([[2.0, 0.3, 0.0],
                  [-0.4, 1.5, 0.0],
                  [0.25, 0.1, 1.2]])


@pytest.mark.parametrize("fd", [P.PADE6, P.PENTA10])
@pytest.mark.parametrize("bc", [(PER, PER, PER), (PER, DIR, DIR), (DIR, DIR, PER)])
def test_affine_map_metrics_are_exact(fd, bc):
    N = (16, 18, 20)
    X, trans = affine(A_GEN, N, bc)
    out = P.metrics((fd, fd, fd), bc, *X, trans)
    Ainv = np.linalg.inv(A_GEN)
    # exact up to rounding; J comes through three nested derivatives (x ->
    # V -> J_ij / J -> x J_i), each of which amplifies rounding noise by about
    # 1 / d(xi), and the one-sided closures more so: a few 1e-10 at worst.  An
    # index, sign or transposition mistake is O(1).
    tol = 1e-12 if bc == (PER, PER, PER) else 1e-9
    assert np.max(np.abs(out[0] - 1.0 / np.linalg.det(A_GEN))) < tol
    for i in range(3):
        for j in range(3):
            assert np.max(np.abs(out[1 + 3 * i + j] - Ainv[i, j])) < tol, (i, j)


def test_cd2_on_a_periodic_direction_is_rejected():
    X, trans = affine(A_GEN, (12, 12, 12), (PER, PER, PER))
    with pytest.raises(RuntimeError):
        P.metrics((P.CD2, P.CD2, P.CD2), (PER, PER, PER), *X, trans)


def _smooth(N, bc, amp):
    xi, eta, zeta = grid(N, bc)
    t = np.pi
    X = [xi + amp * np.sin(t * eta) * np.sin(t * zeta),
         eta + amp * np.sin(t * zeta) * np.sin(t * xi),
         zeta + amp * np.sin(t * xi) * np.sin(t * eta)]
    D = np.zeros((3, 3) + xi.shape)
    D[0, 0] = D[1, 1] = D[2, 2] = 1.0
    D[0, 1] = amp * t * np.cos(t * eta) * np.sin(t * zeta)
    D[0, 2] = amp * t * np.sin(t * eta) * np.cos(t * zeta)
    D[1, 2] = amp * t * np.cos(t * zeta) * np.sin(t * xi)
    D[1, 0] = amp * t * np.sin(t * zeta) * np.cos(t * xi)
    D[2, 0] = amp * t * np.cos(t * xi) * np.sin(t * eta)
    D[2, 1] = amp * t * np.sin(t * xi) * np.cos(t * eta)
    Dm = np.moveaxis(D, (0, 1), (-2, -1))
    return X, np.linalg.inv(Dm), np.linalg.det(Dm)


@pytest.mark.parametrize("fd,tol", [(P.PENTA10, 2e-6), (P.PADE6, 2e-4)])
def test_smooth_dirichlet_map_matches_the_analytic_metrics(fd, tol):
    """x = xi + 0.1 sin(pi eta) sin(pi zeta) and cyclic permutations on
    [0, 1]^3 with boundary closures in every direction: J and J_ij against
    the analytic inverse Jacobian, to the scheme's accuracy at 40 points, and
    the error shrinks at the scheme's order under refinement."""
    errs = []
    for n in (20, 40):
        N = (n, n, n)
        X, inv, det = _smooth(N, (DIR, DIR, DIR), 0.1)
        out = P.metrics((fd,) * 3, (DIR, DIR, DIR), *X, [[0] * 3] * 3)
        e = max(np.max(np.abs(out[1 + 3 * i + j] - inv[..., i, j])) for i in range(3) for j in range(3))
        errs.append(max(e, np.max(np.abs(out[0] * det - 1.0))))
    assert errs[1] < tol
    assert errs[0] / errs[1] > 4.0  # at least second order with the one-sided closures


def test_periodic_path_is_exact_only_for_periodic_products():
    """The reference's transformed-periodic derivative keeps the cyclic
    left-hand side, which assumes a periodic derivative.  Products such as
    x_eta z differentiated across a periodic boundary where z jumps and x_eta
    varies do not have one, so on a smooth periodic map the reference's metrics
    carry an error confined to the planes next to the boundary (reproduced
    here; the HIT box and affine maps are exact, above)."""
    n = 32
    N = (n, n, n)
    bc = (PER, PER, PER)
    xi, eta, zeta = grid(N, bc)
    t = 2 * np.pi
    X = [xi + 0.05 * np.sin(t * eta) * np.sin(t * zeta),
         eta + 0.05 * np.sin(t * zeta) * np.sin(t * xi),
         zeta + 0.05 * np.sin(t * xi) * np.sin(t * eta)]
    out = P.metrics((P.PADE6,) * 3, bc, *X, [[1, 0, 0], [0, 1, 0], [0, 0, 1]])
    D = np.zeros((3, 3) + xi.shape)
    D[0, 0] = D[1, 1] = D[2, 2] = 1.0
    D[0, 1] = 0.05 * t * np.cos(t * eta) * np.sin(t * zeta)
    D[0, 2] = 0.05 * t * np.sin(t * eta) * np.cos(t * zeta)
    D[1, 2] = 0.05 * t * np.cos(t * zeta) * np.sin(t * xi)
    D[1, 0] = 0.05 * t * np.sin(t * zeta) * np.cos(t * xi)
    D[2, 0] = 0.05 * t * np.cos(t * xi) * np.sin(t * eta)
    D[2, 1] = 0.05 * t * np.sin(t * xi) * np.cos(t * eta)
    Dm = np.moveaxis(D, (0, 1), (-2, -1))
    cof = np.linalg.inv(Dm) * np.linalg.det(Dm)[..., None, None]
    # J12 / J = x_zeta z_eta - x_eta z_zeta comes from the zeta derivative of
    # x_eta z: its error is a boundary layer in zeta, decaying geometrically
    err = np.max(np.abs(out[2] / out[0] - cof[..., 0, 1]), axis=(1, 2))  # per zeta plane
    assert err[0] > 1e-1 and err[-1] > 1e-1
    assert np.max(err[n // 2 - 4:n // 2 + 4]) < 1e-4 * err[0]
    assert np.all(err[1:8] < err[0:7])  # decays away from the boundary


def test_seven_point_typos_act_where_the_reference_has_them():
    """With PENTA10 the xi derivative of Vc3 = y_zeta z takes Nm2c3 in the
    Nm3 slot: an affine map with y_zeta != 0 and z depending on xi loses
    exactness in J21 at the start of every xi line.  The same map with Pade6
    (5-point, no Nm3 slot), and a map with y_zeta = 0 under PENTA10, stay
    exact.  (The other typo, zeta derivative of y, is constant along xi and
    eta for an affine map and so does not reach J there.)"""
    N = (16, 16, 16)
    bc = (PER, PER, PER)
    A = np.array([[1.0, 0.0, 0.0], [0.0, 1.3, 0.4], [0.3, 0.0, 0.9]])
    X, trans = affine(A, N, bc)
    Ainv = np.linalg.inv(A)
    bad = P.metrics((P.PENTA10,) * 3, bc, *X, trans)
    good = P.metrics((P.PADE6,) * 3, bc, *X, trans)
    err = lambda out: max(np.max(np.abs(out[1 + 3 * i + j] - Ainv[i, j])) for i in range(3) for j in range(3))
    assert err(good) < 1e-10
    assert err(bad) > 1e-6
    A2 = np.array([[1.0, 0.0, 0.3], [0.0, 1.3, 0.0], [0.2, 0.0, 0.9]])
    X2, trans2 = affine(A2, N, bc)
    ok = P.metrics((P.PENTA10,) * 3, bc, *X2, trans2)
    Ainv2 = np.linalg.inv(A2)
    assert max(np.max(np.abs(ok[1 + 3 * i + j] - Ainv2[i, j])) for i in range(3) for j in range(3)) < 1e-10


def test_periodic_translations_follow_the_reference_face_checks():
    disp = [[6.0, 0.5, 0.0], [0.0, 6.0, 0.0], [0.0, 0.0, 6.0]]
    import ctypes
    L = P.lib()
    L.orc_periodic_translations.restype = None
    I3 = ctypes.c_int * 3
    D9 = ctypes.c_double * 9

    def run(bc, lo, hi):
        tr, st = D9(), I3()
        L.orc_periodic_translations(I3(*bc), I3(*lo), I3(*hi), D9(*sum(disp, [])), tr, st)
        return np.array(tr[:]).reshape(3, 3), list(st)

    tr, st = run((PER, PER, PER), (0, 0, 0), (0, 0, 0))
    assert st == [1, 1, 1] and np.array_equal(tr, np.array(disp))
    # x periodic, y faces not PERIODIC: the x translation is not set (the
    # reference reads the y faces for x), unless x is INTERNALLY_PERIODIC
    tr, st = run((PER, DIR, PER), (0, 2, 0), (0, 2, 0))
    assert st[0] == 0
    tr, st = run((PER, DIR, PER), (1, 2, 0), (1, 2, 0))
    assert st[0] == 1 and not tr[0].any()
