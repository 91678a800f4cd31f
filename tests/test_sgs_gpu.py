"""The SGS models and the CFL time step over caller-owned fields, through the
C ABI (mc3d_sgs_*, mc3d_cfl_dt), against the oracle.

The oracle evaluates the model inside its RHS probe (orc_solver_rhs: the
reference's preStepDerivatives at the initial state).  The handle gets the
same Cartesian gradients the oracle forms (the oracle's own line derivatives,
gradU = J^T d/dxi in the reference's order, src/CurvilinearCSolver_Core.cpp:548-561),
in the x-fastest layout and in the reference's Y and Z pencils.

Tolerances (normwise relative, max|gpu - oracle| / max|oracle|): Vreman
1e-12 (the pointwise formula is evaluated bitwise alike; LOCAL / averaging
adds line solves within 1e-12); DSM 1e-10 for mu_sgs and tau_kk and 1e-9 for
k_sgs, C, C_I (a chain of about 40 test filters and ratios of filtered
contractions); Pr_t 1e-7 under XI averaging, where it divides by a small
line mean (measured amplification of a 1e-15 input perturbation: 4.7e-9).
CFL: dt bitwise equal (a maximum of the same per-point expression).
"""
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

XP, YP, ZP = 0, 1, 2
R_GAS = (1.0 / 1.4) / 1.0 / 1.0


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def to_layout(f, layout):
    if layout == XP:
        return np.ascontiguousarray(f)
    if layout == YP:   # [i, k, j]
        return np.ascontiguousarray(f.transpose(2, 0, 1))
    return np.ascontiguousarray(f.transpose(1, 2, 0))  # [j, i, k]


def from_layout(g, layout):
    if layout == XP:
        return g
    if layout == YP:
        return np.ascontiguousarray(g.transpose(1, 2, 0))
    return np.ascontiguousarray(g.transpose(2, 0, 1))


def sheared_metrics(N):
    """Constant metrics of an affine map x = A xi (A non-symmetric), so every
    J_ij enters the Cartesian gradient: J_ij = (A^-1)_ij, J = 1/det A."""
    A = 2 * np.pi * np.array([[1.0, 0.12, -0.05], [0.07, 1.1, 0.09], [-0.04, 0.03, 0.95]])
    Ai = np.linalg.inv(A)
    one = np.ones((N, N, N))
    return [one / np.linalg.det(A)] + [Ai[i, j] * one for i in range(3) for j in range(3)]


def cart_grad(met, d):
    """g[j] = J1j d0 + J2j d1 + J3j d2 (the reference's order)."""
    J = [[met[1 + 3 * k + j] for j in range(3)] for k in range(3)]  # J[k][j] = J_{k+1, j+1}
    return [J[0][j] * d[0] + J[1][j] * d[1] + J[2][j] * d[2] for j in range(3)]


def thermal_ic(ic, amp):
    """The HIT field with smooth density and pressure fluctuations of relative
    size amp.  The HIT initial field has uniform rho and p, so its temperature
    is uniform and the DSM's K_i = (rho u_i T)^ - (rho u_i)^ (rho T)^ / rho and
    N_i = rho |S| dT/dx_i vanish identically: without tau_kk (which adds
    temperature fluctuations) k_sgs is then a ratio of rounding residues."""
    n = ic["rho"].shape[0]
    t = 2 * np.pi * np.arange(n) / n
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    out = dict(ic)
    out["rho"] = ic["rho"] * (1 + amp * np.sin(x + 2 * y) * np.cos(z))
    out["p"] = ic["p"] * (1 + amp * np.cos(2 * x - y + z))
    return out


def setup(N, les, les_avg, seed, k0, Mt, fd="PENTA10", thermal=0.0, **extra):
    case = dict(fd=fd, filt="COMPACT8", alpha_f=0.49, rk="KENRK4", les=les, les_avg=les_avg, mu_ref=1e-4,
                **extra)
    ic = hit.hit_initial_condition(N, k0=k0, Mt=Mt, seed=seed)
    if thermal:
        ic = thermal_ic(ic, thermal)
    met = sheared_metrics(N)
    o = P.Solver(hit.oracle_config(N, **case), met, ic)
    st = {k: o.field(k) for k in ("rho", "rhoU", "rhoV", "rhoW", "U", "V", "W", "p", "T")}
    o.rhs()
    dd = 1.0 / N
    sch = {"PENTA10": P.PENTA10, "PADE6": P.PADE6}[fd]
    d = {q: [P.deriv_field(sch, k, P.PERIODIC, st[q], dd) for k in range(3)] for q in ("U", "V", "W")}
    gradU = cart_grad(met, d["U"]) + cart_grad(met, d["V"]) + cart_grad(met, d["W"])
    return case, o, st, met, gradU, sch, dd


def dev(torch, a, layout):
    return torch.from_numpy(to_layout(a, layout)).cuda()


AVGS = ["NONE", "LOCAL", "GLOBAL", "XI1_AVG", "XI2_AVG", "XI3_AVG"]


@pytest.mark.parametrize("les_avg", AVGS)
def test_vreman_in_every_layout(torch_cuda, les_avg):
    """VremanSGS::getSGSViscosity with doAveraging (src/VremanSGS.hpp:43-96,
    src/AbstractSGS.cpp:13-164) over caller fields."""
    from mpi_3dcompact_amd.sgs import Sgs
    torch = torch_cuda
    N = 24
    case, o, st, met, gradU, _, _ = setup(N, "VREMAN", les_avg, seed=7, k0=3.0, Mt=0.3)
    want = o.field("mu_sgs")
    for layout in (XP, YP, ZP):
        _say(f"vreman avg={les_avg} layout={layout}")
        s = Sgs(hit.options_text(N, **case), dev(torch, met[0], layout), layout=layout)
        mu = torch.empty((N, N, N), dtype=torch.float64, device="cuda")
        s.vreman([dev(torch, g, layout) for g in gradU], dev(torch, st["rho"], layout), mu)
        torch.cuda.synchronize()
        got = from_layout(mu.cpu().numpy(), layout)
        e = relerr(got, want)
        _say(f"  mu_sgs {e:.2e}")
        assert e <= 1e-12, (les_avg, layout, e)
        s.close()


DSM_CASES = [
    # (averaging, k0, M_t, test filter, tau_kk, {field: tolerance})
    ("GLOBAL", 3.0, 0.3, 1, True, {"mu_sgs": 1e-10, "taukk": 1e-10, "k_sgs": 1e-9, "C": 1e-9, "CI": 1e-9,
                                   "Prt": 1e-9}),
    ("XI1_AVG", 3.0, 0.3, 2, True, {"mu_sgs": 1e-10, "taukk": 1e-10, "k_sgs": 1e-9, "C": 1e-9, "CI": 1e-9,
                                    "Prt": 1e-7}),
    ("XI2_AVG", 3.0, 0.3, 1, False, {"mu_sgs": 1e-10, "k_sgs": 1e-9, "C": 1e-9, "Prt": 1e-7}),
    # LOCAL divides by point-wise filtered contractions that come near zero;
    # on a gentle field mu_sgs and tau_kk are well conditioned
    ("LOCAL", 1.5, 0.05, 1, True, {"mu_sgs": 1e-10, "taukk": 1e-10}),
]


@pytest.mark.parametrize("case", DSM_CASES, ids=[c[0] + f"-filter{c[3]}" for c in DSM_CASES])
def test_dsm_in_every_layout(torch_cuda, case):
    """DSMSGS::calcMuSGSTaukk then calcKSGS (src/DSMSGS.hpp:124-444) over
    caller fields; the temperature gradient is taken after the tau_kk
    pressure update, as the solver does (src/CurvilinearCSolver_Core.cpp:571-594)."""
    from mpi_3dcompact_amd.sgs import Sgs
    torch = torch_cuda
    avg, k0, Mt, filt, taukk_on, tols = case
    N = 24
    # without tau_kk the temperature needs fluctuations of its own (thermal_ic)
    c, o, st, met, gradU, sch, dd = setup(N, "DSM", avg, seed=5, k0=k0, Mt=Mt, dsm_filter=filt,
                                          use_taukk=taukk_on, coeff_range=True, thermal=0.0 if taukk_on else 0.05)
    want = {k: o.field(k) for k in tols}
    for layout in (XP, YP, ZP):
        _say(f"dsm avg={avg} filter={filt} taukk={taukk_on} layout={layout}")
        s = Sgs(hit.options_text(N, **c), dev(torch, met[0], layout), layout=layout)
        f = {k: dev(torch, st[k], layout) for k in ("rho", "rhoU", "rhoV", "rhoW")}
        mu = torch.empty_like(f["rho"])
        tk = torch.zeros_like(f["rho"])
        s.dsm_mu_taukk([dev(torch, g, layout) for g in gradU], f["rho"], f["rhoU"], f["rhoV"], f["rhoW"], mu, tk)
        torch.cuda.synchronize()
        taukk = from_layout(tk.cpu().numpy(), layout)
        # p -= tau_kk / 2, T = p / (rho R), then the temperature gradient
        p2 = st["p"] - 0.5 * taukk
        T2 = p2 / (st["rho"] * R_GAS)
        dT = [P.deriv_field(sch, k, P.PERIODIC, T2, dd) for k in range(3)]
        gradT = cart_grad(met, dT)
        ks = torch.empty_like(f["rho"])
        s.dsm_ksgs([dev(torch, g, layout) for g in gradT], f["rho"], f["rhoU"], f["rhoV"], f["rhoW"],
                   dev(torch, T2, layout), ks)
        torch.cuda.synchronize()
        got = {"mu_sgs": from_layout(mu.cpu().numpy(), layout), "taukk": taukk,
               "k_sgs": from_layout(ks.cpu().numpy(), layout)}
        co = {k: torch.empty_like(f["rho"]) for k in ("C", "CI", "Prt")}
        s.coefficients(co["C"], co["CI"], co["Prt"])
        torch.cuda.synchronize()
        for name in ("C", "CI", "Prt"):
            got[name] = from_layout(co[name].cpu().numpy(), layout)
        for name, tol in tols.items():
            e = relerr(got[name], want[name])
            _say(f"  {name} {e:.2e}")
            assert e <= tol, (avg, layout, name, e)
        s.close()


def test_cfl_dt_matches_oracle(torch_cuda):
    """calcDtFromCFL over caller fields equals the oracle's first dt bitwise;
    two stacked boxes each get their own."""
    from mpi_3dcompact_amd.sgs import cfl_dt
    torch = torch_cuda
    N = 24
    met = sheared_metrics(N)
    ics = [hit.hit_initial_condition(N, k0=3.0, Mt=mt, seed=s) for s, mt in ((1, 0.3), (2, 0.1))]
    want = []
    fields = {k: [] for k in ("U", "V", "W", "sos")}
    for ic in ics:
        cfg = hit.oracle_config(N, cfl=0.5)
        o0 = P.Solver(cfg, met, ic)
        for k in fields:
            fields[k].append(o0.field(k))
        o = P.Solver(cfg, met, ic)
        o.steps(1)
        want.append(o.dt)
    cat = {k: torch.from_numpy(np.concatenate([a.ravel() for a in v])).cuda() for k, v in fields.items()}
    J = [torch.from_numpy(np.concatenate([met[i].ravel()] * 2)).cuda() for i in (1, 5, 9)]
    out = cfl_dt(cat["U"], cat["V"], cat["W"], cat["sos"], J[0], J[1], J[2], (1.0 / N,) * 3, 0.5, members=2)
    got = out.cpu().numpy()
    _say(f"cfl dt {got[:, 0]} oracle {want}")
    assert got[0, 0] == want[0] and got[1, 0] == want[1]
    assert np.allclose(got[:, 0] * got[:, 1], 0.5, rtol=1e-15)
    # one box alone gives the same as its part of the stacked call
    one = cfl_dt(cat["U"][:N ** 3], cat["V"][:N ** 3], cat["W"][:N ** 3], cat["sos"][:N ** 3],
                 J[0][:N ** 3], J[1][:N ** 3], J[2][:N ** 3], (1.0 / N,) * 3, 0.5)
    assert one.cpu().numpy()[0, 0] == want[0]


def _scaled_closure(seed, lo=-1e30, hi=1e30):
    from mpi_3dcompact_amd.closure import Closure
    from oracle import orc_closure as C
    p = C.init_params(seed)
    p[-2:] = [0.25, -0.1]
    c = Closure(seed)
    c.set_params(p)
    im = np.array([0.0, 0.2, 0.0, -0.01, 0.0, 1.0, -3.0], np.float32)
    isd = np.array([0.1, 0.5, 0.2, 0.1, 0.05, 1.0, 0.5], np.float32)
    om, osd = np.array([0.01, 0.005], np.float32), np.array([0.02, 0.01], np.float32)
    c.set_scaling(im, isd, om, osd, np.full(7, lo), np.full(7, hi))
    return c


@pytest.mark.parametrize("les_avg,lo,hi", [("NONE", -1e30, 1e30), ("XI1_AVG", -1e30, 1e30),
                                           ("LOCAL", -0.5, 0.5)],
                         ids=["none", "xi1", "local-fallback"])
def test_learned_matches_the_solver_closure(torch_cuda, les_avg, lo, hi):
    """The learned closure over caller fields gives what the solver's
    LESMODEL LEARNED gives on the same state (the solver's closure is pinned to
    the oracle in test_closure_gpu.py): the network, the feature averaging
    along the layout's axes and, with a narrow trusted range, the DSM fallback
    on the stages whose points left it.  The state has temperature
    fluctuations (thermal_ic): the fallback DSM runs without tau_kk, and on the
    plain HIT field its k_sgs would be a ratio of rounding residues.  Under
    LOCAL averaging in another layout the line solves round differently and
    the DSM divides by point-wise filtered contractions: the oracle moves by
    2.5e-11 (mu_sgs) and 2.9e-9 (k_sgs) under a 1e-15 relative perturbation of
    this state, so that case is held to 1e-9 / 1e-7; every other case to 1e-12."""
    from mpi_3dcompact_amd.sgs import Sgs
    from mpi_3dcompact_amd.solver import Solver
    torch = torch_cuda
    N = 32
    opts = hit.options_text(N, fd="PADE6", filt="COMPACT10", rk="TVDRK3", les="LEARNED", les_avg=les_avg,
                            closure_fallback=lo > -1e30)
    c = _scaled_closure(31, lo, hi)
    met = hit.identity_metrics(N)
    g = Solver(opts, (N, N, N))
    g.set_metrics(met)
    g.set_initial(thermal_ic(hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=6), 0.05))
    g.set_closure(c)
    g.eval_sgs()
    d = {q: [g.field(f"d{q}{k}") for k in range(3)] for q in "UVWT"}
    gradU = cart_grad(met, d["U"]) + cart_grad(met, d["V"]) + cart_grad(met, d["W"])
    gradT = cart_grad(met, d["T"])
    want = {k: g.field(k) for k in ("mu_sgs", "k_sgs")}
    state = {k: g.field(k) for k in ("rho", "rhoU", "rhoV", "rhoW", "T")}
    for layout in (XP, YP):
        _say(f"learned avg={les_avg} range=[{lo}, {hi}] layout={layout}")
        s = Sgs(opts, dev(torch, met[0], layout), layout=layout)
        s.set_closure(c)
        f = {k: dev(torch, v, layout) for k, v in state.items()}
        mu = torch.empty_like(f["rho"])
        ks = torch.empty_like(f["rho"])
        s.learned([dev(torch, x, layout) for x in gradU], f["rho"], mu, ks,
                  gradT=[dev(torch, x, layout) for x in gradT], rhoU=f["rhoU"], rhoV=f["rhoV"], rhoW=f["rhoW"],
                  T=f["T"])
        torch.cuda.synchronize()
        ev, outside, fb_stages, _ = s.closure_stats()
        _say(f"  stats evaluations {ev} outside {outside} fallback stages {fb_stages}")
        assert ev == 1 and (fb_stages == 1) == (lo > -1e30)
        for name, t in (("mu_sgs", mu), ("k_sgs", ks)):
            e = relerr(from_layout(t.cpu().numpy(), layout), want[name])
            tol = 1e-12
            if les_avg == "LOCAL" and layout != XP:
                tol = {"mu_sgs": 1e-9, "k_sgs": 1e-7}[name]
            _say(f"  {name} {e:.2e} (tolerance {tol:.0e})")
            assert e <= tol, (layout, name, e)
        s.close()
