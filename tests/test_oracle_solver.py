"""Pins of the oracle's solver step (CPU).

* RK: the amplification factor of one step on y' = lambda y, run through the
  oracle's own register-update code, matches exp(z) to the scheme's order.
* RHS: the oracle's J*(-div F) equals an independent spectral evaluation of the
  compressible Navier-Stokes equations (conservative Cartesian form, power-law
  viscosity, Fourier heat flux) on smooth periodic fields, inviscid and viscous.
* A uniform flow is preserved exactly enough; mass, momentum and total energy
  are conserved over several steps (periodic operators have zero column sums and
  the filters preserve the mean).
"""
import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

GAMMA = 1.4


@pytest.mark.parametrize("rk,dual,order", [(P.TVDRK3, 1, 3), (P.RK4, 1, 4), (P.KENRK4, 1, 4),
                                           (P.LSLDDRK4, 1, 4), (P.LSLDDRK4, 2, 4)])
def test_rk_order(rk, dual, order):
    e = [abs(P.rk_amplification(rk, z, dual) - np.exp(z)) for z in (-0.2, -0.1)]
    assert abs(np.log2(e[0] / e[1]) - (order + 1)) < 0.3
    # and the scheme is not of higher order (the error is not at rounding level)
    assert e[1] > 1e-12


def _spectral_rhs(rho, u, v, w, p, mu_ref):
    N = rho.shape[0]
    k = np.fft.fftfreq(N, d=1.0 / N)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")

    def d(f, kk):
        return np.real(np.fft.ifftn(1j * kk * np.fft.fftn(f)))

    dx = lambda f: d(f, kx)
    dy = lambda f: d(f, ky)
    dz = lambda f: d(f, kz)
    R = 1.0 / GAMMA
    cp = R * GAMMA / (GAMMA - 1.0)
    T = p / (rho * R)
    mu = mu_ref * T ** 0.76
    kcond = cp * mu / 0.7
    vel = [u, v, w]
    g = [[dx(q), dy(q), dz(q)] for q in vel]  # g[i][j] = d u_i / d x_j
    divu = g[0][0] + g[1][1] + g[2][2]
    tau = [[mu * (g[i][j] + g[j][i] - (2.0 / 3.0) * divu * (i == j)) for j in range(3)] for i in range(3)]
    gT = [dx(T), dy(T), dz(T)]
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    D = [dx, dy, dz]
    out = [-sum(D[j](rho * vel[j]) for j in range(3))]
    for i in range(3):
        out.append(-sum(D[j](rho * vel[i] * vel[j] + p * (i == j) - tau[i][j]) for j in range(3)))
    out.append(-sum(D[j]((E + p) * vel[j] - sum(vel[i] * tau[i][j] for i in range(3)) - kcond * gT[j])
                    for j in range(3)))
    return out


@pytest.mark.parametrize("mu_ref", [0.0, 0.3])
@pytest.mark.parametrize("scheme", ["PENTA10", "PADE6"])
def test_rhs_matches_spectral_navier_stokes(mu_ref, scheme):
    N = 32
    x = 2 * np.pi * np.arange(N) / N
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rho = 1.0 + 0.1 * np.sin(X) * np.cos(Y)
    u = 0.3 * np.sin(Y + 0.5) + 0.1 * np.cos(Z)
    v = 0.2 * np.cos(Z) - 0.1 * np.sin(X)
    w = 0.1 * np.sin(X + Z)
    p = (1.0 + 0.05 * np.cos(X - Y)) / GAMMA
    cfg = hit.oracle_config(N, fd=scheme, mu_ref=mu_ref)
    s = P.Solver(cfg, hit.identity_metrics(N), dict(rho=rho, U=u, V=v, W=w, p=p))
    got = s.rhs()
    ref = _spectral_rhs(rho, u, v, w, p, mu_ref)
    tol = 2e-5 if scheme == "PENTA10" else 2e-3
    for q in range(5):
        scale = np.max(np.abs(ref[q]))
        assert np.max(np.abs(got[q] - ref[q])) <= tol * scale, (q, np.max(np.abs(got[q] - ref[q])) / scale)


def test_uniform_flow_preserved():
    N = 12
    one = np.ones((N, N, N))
    ic = dict(rho=1.2 * one, U=0.3 * one, V=-0.2 * one, W=0.1 * one, p=one / GAMMA)
    cfg = hit.oracle_config(N, fd="PENTA10", filt="COMPACT8", rk="KENRK4", les="VREMAN",
                            les_avg="LOCAL", mu_ref=1e-3)
    s = P.Solver(cfg, hit.identity_metrics(N), ic)
    s.steps(3)
    for name, val in [("rho", 1.2), ("rhoU", 0.36), ("rhoV", -0.24), ("rhoW", 0.12)]:
        np.testing.assert_allclose(s.field(name), val, rtol=0, atol=1e-13)


@pytest.mark.parametrize("rk,les,avg", [("TVDRK3", "NOMODEL", "NONE"), ("KENRK4", "VREMAN", "LOCAL"),
                                        ("LSLDDRK4", "NOMODEL", "NONE"), ("RK4", "VREMAN", "GLOBAL")])
def test_conservation(rk, les, avg):
    N = 16
    ic = hit.hit_initial_condition(N, k0=3.0, Mt=0.3, seed=1)
    cfg = hit.oracle_config(N, fd="PADE6", filt="COMPACT10", rk=rk, les=les, les_avg=avg, mu_ref=1e-3)
    s = P.Solver(cfg, hit.identity_metrics(N), ic)
    tot0 = {n: s.field(n).sum() for n in ("rho", "rhoU", "rhoV", "rhoW", "rhoE")}
    s.steps(4)
    scale = {"rho": N ** 3, "rhoU": N ** 3 * 0.3, "rhoV": N ** 3 * 0.3, "rhoW": N ** 3 * 0.3,
             "rhoE": N ** 3 * 1.8}
    for n, t0 in tot0.items():
        assert abs(s.field(n).sum() - t0) <= 1e-12 * scale[n], n
    assert s.timestep == 4 and s.dt > 0


def test_hit_initial_condition_properties():
    N = 32
    U, V, W = hit.hit_velocity(N, k0=4.0, Mt=0.3, seed=5)
    # turbulent Mach number and zero mean
    assert abs(np.sqrt(np.mean(U ** 2 + V ** 2 + W ** 2)) - 0.3) < 1e-12
    for f in (U, V, W):
        assert abs(f.mean()) < 1e-14
    # spectrally divergence free
    k = np.fft.fftfreq(N, d=1.0 / N)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    div = kx * np.fft.fftn(U) + ky * np.fft.fftn(V) + kz * np.fft.fftn(W)
    assert np.max(np.abs(div)) < 1e-10 * np.max(np.abs(np.fft.fftn(U)))
    # deterministic for a seed
    U2, _, _ = hit.hit_velocity(N, k0=4.0, Mt=0.3, seed=5)
    np.testing.assert_array_equal(U, U2)
