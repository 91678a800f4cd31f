"""Pins of the oracle's time step (test infrastructure checking test
infrastructure): the CFL time step, the time bookkeeping, the stage sequencing
of the four Runge-Kutta schemes with the momentum-increment accumulation (T1)
and the 2N register aliasing (T4), the filter's round robin (T2), and the
Vreman model's closed forms including its abeta typo (T6).

The step is checked against an independent restatement: textbook Runge-Kutta
recurrences (Shu-Osher TVDRK3, classical RK4, Williamson 2N form for KenRK4
and LSLDDRK4) written over separate arrays (no register reuse), driven by the
oracle's own right-hand side at each stage state (orc_solver_rhs, pinned
against a spectral Navier-Stokes evaluation in test_oracle_solver.py), with
the reference's momentum rule added on top:
    continuity, energy:  K_s = dt R(q_s)
    momentum:            K_s = dt R(q_s) + dt J K_{s-1}    (src/CurvilinearCSolver_Core.cpp:839-840)
with K_{s-1} carried across steps (the k2 registers are never reset).  The
2N coefficients typed here are pinned by the fourth-order conditions of their
Butcher form and by the stage times the reference lists
(src/LSLDDRK4.hpp:40-78, src/KenRK4.hpp:30-40).  Each pin is shown to fail
when the rule it checks is dropped.
"""
import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

GAMMA = 1.4
R_GAS = (1.0 / GAMMA) / 1.0 / 1.0
MOM = (1, 2, 3)

# ---------------------------------------------------------------------------
# Runge-Kutta coefficients (textbook forms)

KEN_A = [0.0, -6234157559845.0 / 12983515589748.0, -6194124222391.0 / 4410992767914.0,
         -31623096876824.0 / 15682348800105.0, -12251185447671.0 / 11596622555746.0]
KEN_B = [494393426753.0 / 4806282396855.0, 4047970641027.0 / 5463924506627.0,
         9795748752853.0 / 13190207949281.0, 4009051133189.0 / 8539092990294.0,
         1348533437543.0 / 7166442652324.0]
LS5_A = [0.0, -0.60512264332862261228, -2.04375640234761394333, -0.74069990637544192841,
         -4.42317651302968168941]
LS5_B = [0.26874543888713438496, 0.8014706973220802933, 0.50515704269422722538, 0.56235680379000296407,
         0.05900655127758823335]
LS5_C = [0.0, 0.26874543888713438496, 0.58522806929524303469, 0.68270664478424678821, 1.1646854837729261436]
LS6_A = [0.0, -0.44127377153877382565, -1.073982008079781868, -1.7063570791256758809,
         -2.7979293162682443056, -4.0913537120919160454]
LS6_B = [0.11584888181285561688, 0.37287699051652864918, 0.73795368921435295698, 0.57981109366311039538,
         1.031284991300145194, 0.15]
LS6_C = [0.0, 0.11584888181285561688, 0.32418503640412806853, 0.61932082035177792368,
         0.80344726663359079059, 0.91841664452065965078]


def butcher_of_2n(A, B):
    """Butcher tableau of the Williamson 2N recurrence dU = A_s dU + K_s,
    u = u + B_s dU (stage s evaluated at the u before its update)."""
    S = len(A)
    a = np.zeros((S + 1, S))
    for s in range(1, S + 1):
        for j in range(1, s + 1):
            tot = 0.0
            for m in range(j, s + 1):
                prod = 1.0
                for l in range(j + 1, m + 1):
                    prod *= A[l - 1]
                tot += B[m - 1] * prod
            a[s, j - 1] = tot
    return a[:S], a[S].copy(), a[:S].sum(axis=1)


def order_conditions(a, b, c):
    """Residuals of the eight conditions of order four."""
    return np.array([b.sum() - 1, b @ c - 1 / 2, b @ c ** 2 - 1 / 3, b @ (a @ c) - 1 / 6, b @ c ** 3 - 1 / 4,
                     b @ (c * (a @ c)) - 1 / 8, b @ (a @ c ** 2) - 1 / 12, b @ (a @ (a @ c)) - 1 / 24])


@pytest.mark.parametrize("A,B,C", [(KEN_A, KEN_B, None), (LS5_A, LS5_B, LS5_C), (LS6_A, LS6_B, LS6_C)],
                         ids=["kenrk4", "lsldd-5", "lsldd-6"])
def test_2n_coefficients_are_fourth_order(A, B, C):
    a, b, c = butcher_of_2n(A, B)
    assert np.max(np.abs(order_conditions(a, b, c))) < 5e-15
    if C is not None:
        assert np.max(np.abs(c - np.array(C))) < 2e-15
    # a one-digit slip in any coefficient breaks an order condition
    for k in range(1, len(A)):
        A2 = list(A)
        A2[k] *= 1.0 + 1e-6
        assert np.max(np.abs(order_conditions(*butcher_of_2n(A2, B)))) > 1e-8


# ---------------------------------------------------------------------------
# the restated step


def smooth_state(N, seed, amp=0.08, noise=0.0):
    """rho, U, V, W, p of a few long waves (|k| <= 2) on the box, plus white
    noise of relative size `noise` (which the filters act on strongly)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(N) / N, np.arange(N) / N, np.arange(N) / N, indexing="ij")

    def wave():
        f = np.zeros((N, N, N))
        for _ in range(3):
            k = rng.integers(-2, 3, size=3)
            f += np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 2 * np.pi))
        return f / 3 + noise * rng.standard_normal((N, N, N))

    return dict(rho=1.0 + 0.3 * amp * wave(), U=amp * wave(), V=amp * wave(), W=amp * wave(),
                p=1.0 / GAMMA + 0.3 * amp * wave())


def primitive(q):
    rho = q[0]
    U, V, W = q[1] / rho, q[2] / rho, q[3] / rho
    p = (GAMMA - 1) * (q[4] - 0.5 * rho * (U * U + V * V + W * W))
    return dict(rho=rho, U=U, V=V, W=W, p=p)


def conserved(ic):
    rho, U, V, W, p = (ic[k] for k in ("rho", "U", "V", "W", "p"))
    return [rho, rho * U, rho * V, rho * W, p / (GAMMA - 1.0) + 0.5 * rho * (U * U + V * V + W * W)]


def rhs_at(cfg, met, q):
    """R = J (-div F) of the five equations at conserved state q (oracle probe)."""
    return P.Solver(cfg, met, primitive(q)).rhs()


def cfl_dt(cfg, met, q, cfl):
    """calcDtFromCFL (src/CurvilinearCSolver_Core.cpp:197-218) written from its definition."""
    ic = primitive(q)
    c = np.sqrt(GAMMA * ic["p"] / ic["rho"])
    speed = (met[1] * (np.abs(ic["U"]) + c) / cfg["dx"] + met[5] * (np.abs(ic["V"]) + c) / cfg["dy"]
             + met[9] * (np.abs(ic["W"]) + c) / cfg["dz"])
    return cfl / speed.max()


def filter_q(q, cfg, order):
    """The five conserved fields filtered along the directions in `order`."""
    kind = P.COMPACT8 if cfg["filter"] == 0 else P.COMPACT10
    out = []
    for f in q:
        for d in order:
            f = P.filter_field(kind, cfg["alphaF"], d, list(cfg["bc"]), f)
        out.append(f)
    return out


ROUND_ROBIN = {1: (1, 2, 0), 2: (2, 0, 1), 0: (0, 1, 2)}  # filterTimeStep % 3 -> Y Z X, Z X Y, X Y Z


class Stepper:
    """Textbook time steps of the reference's schemes over separate arrays."""

    def __init__(self, cfg, met, q0, t1=True, round_robin=True):
        self.cfg, self.met, self.q = cfg, met, [a.copy() for a in q0]
        self.J = met[0]
        self.kprev = [np.zeros_like(q0[0]) for _ in MOM]  # momentum k2 registers, zero at the start
        self.t1, self.round_robin = t1, round_robin
        self.timestep, self.filter_ts, self.dual = 0, 0, 1

    def K(self, qs, dt):
        R = rhs_at(self.cfg, self.met, qs)
        K = [dt * r for r in R]
        if self.t1:
            for n, m in enumerate(MOM):
                K[m] = K[m] + dt * self.J * self.kprev[n]
        self.kprev = [K[m] for m in MOM]
        return K

    def step(self, dt):
        cfg, q0 = self.cfg, self.q
        lin = lambda *terms: [sum(c * t[i] for c, t in terms) for i in range(5)]
        rk = cfg["rk"]
        if rk == 0:  # TVDRK3, Shu & Osher
            K1 = self.K(q0, dt)
            q1 = lin((1, q0), (1, K1))
            K2 = self.K(q1, dt)
            q2 = lin((3 / 4, q0), (1 / 4, q1), (1 / 4, K2))
            K3 = self.K(q2, dt)
            qn = lin((1 / 3, q0), (2 / 3, q2), (2 / 3, K3))
        elif rk == 1:  # classical RK4
            K1 = self.K(q0, dt)
            K2 = self.K(lin((1, q0), (1 / 2, K1)), dt)
            K3 = self.K(lin((1, q0), (1 / 2, K2)), dt)
            K4 = self.K(lin((1, q0), (1, K3)), dt)
            qn = lin((1, q0), (1 / 6, K1), (1 / 3, K2), (1 / 3, K3), (1 / 6, K4))
        else:  # Williamson 2N
            if rk == 2:
                A, B = KEN_A, KEN_B
            else:
                A, B = (LS5_A, LS5_B) if self.dual == 1 else (LS6_A, LS6_B)
                self.dual = 2 if self.dual == 1 else 1
            u = [a.copy() for a in q0]
            dU = [np.zeros_like(a) for a in q0]
            for s in range(len(A)):
                K = self.K(u, dt)
                dU = lin((A[s], dU), (1, K))
                u = lin((1, u), (B[s], dU))
            qn = u
        self.timestep += 1
        if self.timestep % cfg["filter_step"] == 0:
            self.filter_ts += 1
            order = ROUND_ROBIN[self.filter_ts % 3] if self.round_robin else (1, 2, 0)
            qn = filter_q(qn, cfg, order)
        self.q = qn


def rel_incr(a, b, q0):
    """max over fields of |a - b| / |a - q0| (L2): the error relative to the step's change."""
    return max(np.linalg.norm(x - y) / max(np.linalg.norm(x - z), 1e-300) for x, y, z in zip(a, b, q0))


def run_both(case, steps, N=16, seed=1, noise=0.0, **flags):
    cfg = hit.oracle_config(N, **case)
    ic = smooth_state(N, seed, noise=noise)
    met = hit.identity_metrics(N)
    o = P.Solver(cfg, met, ic)
    s = Stepper(cfg, met, conserved(ic), **flags)
    errs = []
    for _ in range(steps):
        q0 = [a.copy() for a in s.q]
        dt = cfl_dt(cfg, met, s.q, case.get("cfl", 0.5))
        o.steps(1)
        assert abs(o.dt - dt) <= 1e-15 * dt, (o.dt, dt)
        s.step(o.dt)
        got = [o.field(k) for k in ("rho", "rhoU", "rhoV", "rhoW", "rhoE")]
        errs.append(rel_incr(got, s.q, q0))
        s.q = got  # continue from the oracle's state (its k2 registers carry on)
    return errs


SCHEMES = ["TVDRK3", "RK4", "KENRK4", "LSLDDRK4"]


@pytest.mark.parametrize("rk", SCHEMES)
def test_step_sequencing(rk):
    """Two steps (LSLDDRK4: its 5- then its 6-stage form; FILTERSTEP 2: the
    second step filters) equal the textbook recurrences with the T1 rule to
    1e-9 of each step's change; without T1 they miss by more than 1e-6."""
    case = dict(fd="PADE6", filt="COMPACT10", alpha_f=0.49, rk=rk, cfl=0.5, mu_ref=1e-3, filter_step=2)
    errs = run_both(case, 2)
    assert max(errs) <= 1e-9, errs
    errs_no_t1 = run_both(case, 2, t1=False)
    assert min(errs_no_t1) > 1e-6, errs_no_t1


def test_filter_round_robin_is_observable():
    """With Dirichlet eta and zeta the directional filters no longer commute
    (pass-through lines, src/Compact8Filter.cpp:148-222), so the round robin of
    filterConservedData (src/CurvilinearCSolver_Core.cpp:990-1134) shows:
    three steps filter Y Z X, Z X Y, X Y Z; a fixed order misses."""
    N = 16
    case = dict(fd="PENTA10", filt="COMPACT8", alpha_f=0.49, rk="TVDRK3", cfl=0.5, mu_ref=1e-3, filter_step=1,
                faces={"Y0": "INLET", "Y1": "INLET", "Z0": "INLET", "Z1": "INLET"})
    errs = run_both(case, 3, noise=0.02)
    assert max(errs) <= 1e-9, errs
    errs_fixed = run_both(case, 3, noise=0.02, round_robin=False)
    assert errs_fixed[0] <= 1e-9 and min(errs_fixed[1:]) > 1e-7, errs_fixed


# ---------------------------------------------------------------------------
# CFL and time bookkeeping (src/CurvilinearCSolver_Core.cpp:189-237)


def test_cfl_closed_form_uniform_flow():
    """Uniform flow, identity metrics of the 2 pi box: dt = CFL / (N/(2 pi)
    sum_i (|u_i| + c)), c = 1 at rho = 1, p = 1/gamma."""
    N = 16
    u = (0.3, -0.2, 0.1)
    ic = dict(rho=np.ones((N, N, N)), U=np.full((N, N, N), u[0]), V=np.full((N, N, N), u[1]),
              W=np.full((N, N, N), u[2]), p=np.full((N, N, N), 1 / GAMMA))
    o = P.Solver(hit.oracle_config(N, cfl=0.5), hit.identity_metrics(N), ic)
    o.steps(1)
    want = 0.5 / (N / (2 * np.pi) * sum(abs(v) + 1.0 for v in u))
    assert abs(o.dt - want) <= 4e-16 * want
    # the maximum over points: a faster point sets the step
    ic["U"][3, 5, 7] = 2.0
    o = P.Solver(hit.oracle_config(N, cfl=0.5), hit.identity_metrics(N), ic)
    o.steps(1)
    want = 0.5 / (N / (2 * np.pi) * ((2.0 + 1.0) + (0.2 + 1.0) + (0.1 + 1.0)))
    assert abs(o.dt - want) <= 4e-16 * want


def test_time_is_not_advanced_on_the_first_step():
    """timeStep == opt->timeStep: time = opt->time, later time += dt (T11)."""
    N = 16
    ic = smooth_state(N, 3)
    o = P.Solver(hit.oracle_config(N, cfl=0.5), hit.identity_metrics(N), ic)
    dts = []
    for _ in range(3):
        o.steps(1)
        dts.append(o.dt)
    assert o.time == pytest.approx(dts[1] + dts[2], rel=1e-15)
    assert dts[0] != dts[1]


# ---------------------------------------------------------------------------
# Vreman (src/VremanSGS.hpp:43-96), closed forms with NONE averaging

C_VREMAN = 2.5 * 0.17 * 0.17


def vreman_mu(N, U=None, V=None, W=None):
    zero = np.zeros((N, N, N))
    ic = dict(rho=np.ones((N, N, N)), U=zero if U is None else U, V=zero if V is None else V,
              W=zero if W is None else W, p=np.full((N, N, N), 1 / GAMMA))
    cfg = hit.oracle_config(N, fd="PADE6", les="VREMAN", les_avg="NONE", mu_ref=1e-3)
    o = P.Solver(cfg, hit.identity_metrics(N), ic)
    o.rhs()
    return o.field("mu_sgs"), cfg


def grad_along(f, d, N):
    """d f / d x_d = J_dd d f / d xi_d with the oracle's Pade6 (identity metrics)."""
    return hit.identity_metrics(N)[1 + 4 * d] * P.deriv_field(P.PADE6, d, P.PERIODIC, f, 1.0 / N)


def profile(N, d, amp=0.3):
    ax = np.arange(N) / N
    shape = [1, 1, 1]
    shape[2 - d] = N
    return np.broadcast_to(amp * np.sin(2 * np.pi * ax).reshape(shape), (N, N, N)).copy()


def vreman_closed(abeta, bbeta, N):
    """c rho sqrt(max(B, eps) / (abeta + eps)) capped at 0.05 Delta^2 sqrt(2 abeta)
    (rho = 1, Delta^3 = dxi deta dzeta / J of the 2 pi box)."""
    d2 = np.cbrt((1.0 / N) ** 3 * (2 * np.pi) ** 3) ** 2
    m = C_VREMAN * np.sqrt(np.maximum(bbeta * d2 * d2, 1e-12) / (abeta + 1e-12))
    return np.minimum(m, 0.05 * d2 * np.sqrt(2 * abeta)), m


def test_vreman_single_gradient_components():
    """One non-zero gradient component g: B_beta = 0, so the eps floor sets
    mu = c rho sqrt(eps / (abeta + eps)) below the upper clip
    0.05 Delta^2 sqrt(2 abeta).  abeta is g^2 for d u/d y, 2 g^2 for d w/d y
    (d2v3 is summed twice, T6) and 0 for d v/d z (d3v2 is missing: the clip
    gives 0 up to the rounding noise of the other derivatives)."""
    N = 16
    g = grad_along(profile(N, 1), 1, N)
    zero = np.zeros_like(g)
    # where the profile's slope vanishes g is rounding noise: absolute 1e-12 of the largest value
    mu, _ = vreman_mu(N, U=profile(N, 1))          # u(y): d2v1 only
    tol = 1e-12 * mu.max()
    assert np.allclose(mu, vreman_closed(g * g, zero, N)[0], rtol=1e-12, atol=tol)
    mu, _ = vreman_mu(N, W=profile(N, 1))          # w(y): d2v3 only
    assert np.allclose(mu, vreman_closed(2 * g * g, zero, N)[0], rtol=1e-12, atol=tol)
    assert not np.allclose(mu, vreman_closed(g * g, zero, N)[0], rtol=1e-3, atol=tol)
    mu_w = mu
    mu, _ = vreman_mu(N, V=profile(N, 2))          # v(z): d3v2 only
    # abeta holds only the rounding noise of the derivatives across the planes
    assert mu.max() <= 1e-12 * mu_w.max()


def test_vreman_diagonal_gradient():
    """gradU = diag(a, b, c): B_beta = Delta^4 (a^2 b^2 + a^2 c^2 + b^2 c^2),
    abeta = a^2 + b^2 + c^2."""
    N = 16
    U, V, W = profile(N, 0, 0.3), profile(N, 1, 0.2), profile(N, 2, 0.25)
    mu, cfg = vreman_mu(N, U=U, V=V, W=W)
    a, b, c = grad_along(U, 0, N), grad_along(V, 1, N), grad_along(W, 2, N)
    want, m = vreman_closed(a * a + b * b + c * c, a * a * b * b + a * a * c * c + b * b * c * c, N)
    # the 8 points where all three slopes vanish hold rounding noise: absolute 1e-12 of the largest value
    assert np.allclose(mu, want, rtol=1e-12, atol=1e-12 * mu.max())
    assert np.any(want < m) and np.any(want == m)  # both branches of the clip occur
