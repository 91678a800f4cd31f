"""Pins of the CPU oracle against facts the mathematics fixes.

The reference ships no unit tests, golden vectors or fixtures for its line
operators (its only tests exercise the MPI transposes), and it cannot be built
here (its every solver translation unit includes Boost.ProgramOptions, which is
absent).  The oracle is therefore pinned by:
  * dense linear algebra for the banded solvers (numpy.linalg.solve/inv),
  * the closed-form Fourier symbol of each periodic scheme (Lele 1992),
  * polynomial exactness of each closure to its order, and failure one order up,
  * filter transfer-function identities T(0)=1, T(pi)=0 and the order of 1-T,
  * the linear-mapping identity of the transformed-periodic wrap.
Each is chosen so that a dropped term, a wrong sign, a wrong index or a swapped
band fails it.
"""
import numpy as np
import pytest

from oracle import pyoracle as P

RNG = np.random.default_rng(1234)


# --------------------------------------------------------------------------
# banded solvers vs dense solves
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n", [11, 16, 64])
def test_solve_tri_dense(orc, n):
    a, c = RNG.uniform(-1, 1, n), RNG.uniform(-1, 1, n)
    b = 4.0 + RNG.uniform(0, 1, n)
    d = RNG.normal(size=n)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)  # a indexed by row
    np.testing.assert_allclose(orc.solve_tri(a, b, c, d), np.linalg.solve(A, d), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [11, 16, 64])
def test_cyclic_tri_dense(orc, n):
    a, c = RNG.uniform(-1, 1, n), RNG.uniform(-1, 1, n)
    b = 4.0 + RNG.uniform(0, 1, n)
    alpha, beta = 0.7, -0.3
    r = RNG.normal(size=n)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    A[n - 1, 0] = alpha  # bottom-left corner
    A[0, n - 1] = beta   # top-right corner
    np.testing.assert_allclose(orc.cyclic_tri(a, b, c, alpha, beta, r), np.linalg.solve(A, r),
                               rtol=0, atol=1e-13)


def _penta_dense(e, a, d, c, f):
    n = d.size
    A = np.diag(d) + np.diag(c[:-1], 1) + np.diag(f[:-2], 2)
    A += np.diag(a[:-1], -1) + np.diag(e[:-2], -2)  # lower bands indexed by column
    return A


@pytest.mark.parametrize("n", [11, 16, 64])
def test_solve_penta_dense(orc, n):
    e, a, c, f = (RNG.uniform(-1, 1, n) for _ in range(4))
    d = 6.0 + RNG.uniform(0, 1, n)
    b = RNG.normal(size=n)
    A = _penta_dense(e, a, d, c, f)
    np.testing.assert_allclose(orc.solve_penta(e, a, d, c, f, b), np.linalg.solve(A, b),
                               rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [11, 16, 64])
def test_cyclic_penta_dense(orc, n):
    e, a, c, f = (RNG.uniform(-1, 1, n) for _ in range(4))
    d = 6.0 + RNG.uniform(0, 1, n)
    cp = RNG.uniform(-1, 1, 6)
    rhs = RNG.normal(size=n)
    A = _penta_dense(e, a, d, c, f)
    A[0, n - 2], A[0, n - 1], A[1, n - 1] = cp[0], cp[1], cp[2]
    A[n - 2, 0], A[n - 1, 0], A[n - 1, 1] = cp[3], cp[4], cp[5]
    # argument order of the Woodbury routine: (lower2, lower, diag, upper, upper2, rhs, cp)
    x = orc.cyclic_penta(e, a, d, c, f, rhs, cp)
    np.testing.assert_allclose(x, np.linalg.solve(A, rhs), rtol=0, atol=1e-12)


def test_invert4(orc):
    m = RNG.normal(size=(4, 4)) + 3 * np.eye(4)
    np.testing.assert_allclose(orc.invert4(m), np.linalg.inv(m), rtol=0, atol=1e-13)


# --------------------------------------------------------------------------
# periodic derivative schemes: exact on every Fourier mode with k'(theta)
# --------------------------------------------------------------------------
SCHEMES = {
    # alpha, beta, a, b, c  (Lele 1992 family; Pade6 and the pentadiagonal 10th order)
    P.PADE6: (1 / 3, 0.0, 14 / 9, 1 / 9, 0.0),
    P.PENTA10: (1 / 2, 1 / 20, 17 / 12, 101 / 150, 1 / 100),
}


def modified_wavenumber(scheme, th):
    al, be, a, b, c = SCHEMES[scheme]
    return (a * np.sin(th) + b / 2 * np.sin(2 * th) + c / 3 * np.sin(3 * th)) / (
        1 + 2 * al * np.cos(th) + 2 * be * np.cos(2 * th))


@pytest.mark.parametrize("scheme", [P.PADE6, P.PENTA10])
@pytest.mark.parametrize("n", [16, 64, 256])
def test_periodic_symbol(orc, scheme, n):
    h = 1.0 / n
    j = np.arange(n)
    for m in [1, 3, n // 4, n // 2 - 1]:
        th = 2 * np.pi * m / n
        ph = RNG.uniform(0, 2 * np.pi)
        f = np.sin(th * j + ph)
        exact = modified_wavenumber(scheme, th) / h * np.cos(th * j + ph)
        got = orc.deriv_line(scheme, P.PERIODIC, f, h)
        np.testing.assert_allclose(got, exact, rtol=0, atol=2e-13 * n)


@pytest.mark.parametrize("scheme", [P.PADE6, P.PENTA10])
def test_periodic_order(orc, scheme):
    # 1 - k'(th)/th = O(th^6) for Pade6 and O(th^10) for Penta10
    order = 6 if scheme == P.PADE6 else 10
    th1, th2 = 0.8, 0.4
    e1 = 1 - modified_wavenumber(scheme, th1) / th1
    e2 = 1 - modified_wavenumber(scheme, th2) / th2
    assert abs(np.log2(e1 / e2) - order) < 0.35
    # and the oracle realises the same symbol on the grid
    n = 64
    j = np.arange(n)
    th = 2 * np.pi * 8 / n
    got = orc.deriv_line(scheme, P.PERIODIC, np.sin(th * j), 1.0)
    np.testing.assert_allclose(got, modified_wavenumber(scheme, th) * np.cos(th * j), atol=1e-12)


# --------------------------------------------------------------------------
# Dirichlet closures: polynomial exactness to the closure order, and not beyond
# --------------------------------------------------------------------------
def _poly_err(orc, scheme, n, deg):
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    f = (x - 0.37) ** deg
    exact = deg * (x - 0.37) ** (deg - 1) if deg else np.zeros(n)
    got = orc.deriv_line(scheme, P.DIRICHLET, f, h)
    return np.max(np.abs(got - exact)) / max(1.0, np.max(np.abs(exact)))


@pytest.mark.parametrize("scheme,exact_deg", [(P.PADE6, 3), (P.PENTA10, 6)])
def test_dirichlet_polynomial_exactness(orc, scheme, exact_deg):
    n = 40
    for deg in range(exact_deg + 1):
        assert _poly_err(orc, scheme, n, deg) < 1e-11, deg
    assert _poly_err(orc, scheme, n, exact_deg + 1) > 1e-8


def test_dirichlet_interior_matches_periodic(orc):
    # away from the ends the Dirichlet operator must approach the periodic one
    n = 128
    x = np.arange(n) / (n - 1)
    f = np.exp(-((x - 0.5) / 0.1) ** 2)
    for s in (P.PADE6, P.PENTA10):
        d = orc.deriv_line(s, P.DIRICHLET, f, 1 / (n - 1))
        exact = -2 * (x - 0.5) / 0.01 * f
        assert np.max(np.abs(d[40:-40] - exact[40:-40])) < 1e-5 * np.max(np.abs(exact))


def test_cd2_sign_convention(orc):
    """CD2 evaluates (a*f[i+1] + c*f[i-1])/h with a=-1/2, c=+1/2: the interior
    carries the negated central difference (exact for quadratics up to sign);
    the one-sided first row and the wrap rows of the periodic form carry the
    usual sign."""
    n, h = 20, 0.05
    x = np.arange(n) * h
    f = 1.0 + 2.0 * x + 3.0 * x ** 2
    fp = 2.0 + 6.0 * x
    d = orc.deriv_line(P.CD2, P.DIRICHLET, f, h)
    np.testing.assert_allclose(d[1:-1], -fp[1:-1], rtol=1e-12)
    np.testing.assert_allclose(d[0], fp[0], rtol=1e-12)
    np.testing.assert_allclose(d[-1], -fp[-1], rtol=1e-12)
    th = 2 * np.pi / n
    g = np.sin(th * np.arange(n))
    dp = orc.deriv_line(P.CD2, P.PERIODIC, g, 1.0)
    k = np.sin(th)  # central-difference symbol
    np.testing.assert_allclose(dp[1:-1], -k * np.cos(th * np.arange(n))[1:-1], atol=1e-13)
    np.testing.assert_allclose(dp[[0, -1]], k * np.cos(th * np.arange(n))[[0, -1]], atol=1e-13)


# --------------------------------------------------------------------------
# transformed-periodic wrap values
# --------------------------------------------------------------------------
@pytest.mark.parametrize("scheme", [P.PADE6, P.PENTA10, P.CD2])
def test_tpb_linear_map(orc, scheme):
    n = 32
    L = 2 * np.pi
    x = L * np.arange(n) / n
    g = 0.1 * np.sin(2 * np.pi * np.arange(n) / n)
    f = x + g
    wrap = np.array([f[n - 3] - L, f[n - 2] - L, f[n - 1] - L, f[0] + L, f[1] + L, f[2] + L])
    d_tpb = orc.deriv_line(scheme, P.PERIODIC, f, 1.0 / n, tpb=wrap)
    d_g = orc.deriv_line(scheme, P.PERIODIC, g, 1.0 / n)
    if scheme == P.CD2:
        # interior sign of CD2 (see test_cd2_sign_convention) flips the slope there
        np.testing.assert_allclose(d_tpb[[0, -1]], L + d_g[[0, -1]], atol=1e-12)
        np.testing.assert_allclose(d_tpb[1:-1], -L + d_g[1:-1], atol=1e-12)
    else:
        np.testing.assert_allclose(d_tpb, L + d_g, atol=1e-11)


def test_neumann_extrapolation(orc):
    # f[0] is returned so that the one-sided 6-point derivative at the wall is 0:
    # exact for polynomials with zero slope at the wall up to degree 6
    # (six samples plus the slope condition fix seven coefficients).
    xs = np.arange(1, 7, dtype=float)
    for deg in range(2, 7):
        g = lambda x: 1.0 + 0.3 * x ** deg
        assert abs(orc.neumann(P.PADE6, g(xs)) - g(0.0)) < 1e-11 * g(xs).max()
    g = lambda x: 1.0 + 0.3 * x ** 7
    assert abs(orc.neumann(P.PADE6, g(xs)) - g(0.0)) > 1e-3


# --------------------------------------------------------------------------
# filters
# --------------------------------------------------------------------------
def filter_symbol(kind, af, th):
    if kind == P.COMPACT8:
        a = [(93 + 70 * af) / 128, (7 + 18 * af) / 16, (-7 + 14 * af) / 32, (1 - 2 * af) / 16,
             (-1 + 2 * af) / 128]
        num = a[0] + sum(a[k] * np.cos(k * th) for k in range(1, 5))
        return num / (1 + 2 * af * np.cos(th))
    if kind == P.COMPACT10:
        a = [(193 + 126 * af) / 256, (105 + 302 * af) / 256, (-15 + 30 * af) / 64,
             (45 - 90 * af) / 512, (-5 + 10 * af) / 256, (1 - 2 * af) / 512]
        num = a[0] + sum(a[k] * np.cos(k * th) for k in range(1, 6))
        return num / (1 + 2 * af * np.cos(th))
    if kind == P.COMMN5PADE:
        w = [7 / 24, 175 / 768, 5 / 48, 35 / 1536, 0.0, -1 / 1536]
        num = w[0] + 2 * sum(w[k] * np.cos(k * th) for k in range(1, 6))
        return num / (7 / 12 + 2 * 5 / 24 * np.cos(2 * th))
    w = [11 / 16, 15 / 64, -3 / 32, 1 / 64]
    return w[0] + 2 * sum(w[k] * np.cos(k * th) for k in range(1, 4))


FILTERS = [(P.COMPACT8, 0.49), (P.COMPACT8, 0.3), (P.COMPACT10, 0.49), (P.COMPACT10, -0.49),
           (P.COMMN5PADE, 0.0), (P.COMMN5EXP, 0.0)]


@pytest.mark.parametrize("kind,af", FILTERS)
def test_filter_symbol_identities(kind, af):
    assert abs(filter_symbol(kind, af, 0.0) - 1.0) < 1e-14   # mean preserved
    assert abs(filter_symbol(kind, af, np.pi)) < 1e-14        # Nyquist removed


@pytest.mark.parametrize("kind,order", [(P.COMPACT8, 8), (P.COMPACT10, 10)])
def test_filter_order(kind, order):
    e1 = 1 - filter_symbol(kind, 0.49, 0.4)
    e2 = 1 - filter_symbol(kind, 0.49, 0.2)
    assert abs(np.log2(e1 / e2) - order) < 0.3


@pytest.mark.parametrize("kind,af", FILTERS)
@pytest.mark.parametrize("n", [16, 64])
def test_filter_periodic_modes(orc, kind, af, n):
    j = np.arange(n)
    for m in [0, 1, 5, n // 4, n // 2]:
        th = 2 * np.pi * m / n
        f = np.cos(th * j + 0.3)
        got = orc.filter_line(kind, P.PERIODIC, f, af)
        np.testing.assert_allclose(got, filter_symbol(kind, af, th) * f, atol=1e-13)


@pytest.mark.parametrize("kind,af,deg", [(P.COMPACT8, 0.49, 7), (P.COMPACT10, 0.49, 9),
                                         (P.COMMN5PADE, 0.0, 3), (P.COMMN5EXP, 0.0, 3)])
def test_filter_dirichlet_preserves_polynomials(orc, kind, af, deg):
    n = 48
    x = np.linspace(-1, 1, n)
    for d in range(deg + 1):
        f = x ** d
        got = orc.filter_line(kind, P.DIRICHLET, f, af)
        np.testing.assert_allclose(got, f, atol=1e-11)
    # a plausible boundary mistake would break exactness; check the ends filter at all
    g = np.cos(np.pi * np.arange(n))  # Nyquist mode: strongly damped in the interior
    out = orc.filter_line(kind, P.DIRICHLET, g, af)
    assert out[0] == g[0] or kind in (P.COMMN5PADE, P.COMMN5EXP)
    assert np.max(np.abs(out - g)) > 1e-3


def test_filter_passflag(orc):
    # a Dirichlet face in y leaves the x-lines at j=0 and j=Ny-1 untouched
    Nz, Ny, Nx = 12, 14, 16
    f = RNG.normal(size=(Nz, Ny, Nx))
    out = orc.filter_field(P.COMPACT8, 0.49, 0, [P.PERIODIC, P.DIRICHLET, P.PERIODIC], f)
    np.testing.assert_array_equal(out[:, 0, :], f[:, 0, :])
    np.testing.assert_array_equal(out[:, -1, :], f[:, -1, :])
    line = orc.filter_line(P.COMPACT8, P.PERIODIC, f[3, 5, :], 0.49)
    np.testing.assert_array_equal(out[3, 5, :], line)
