"""Ensembles of periodic HIT boxes of uneven sizes: every member of a batched
(stacked, graph-replayed) run against the same member run alone through the
whole-solver ABI, and against the CPU oracle.

Bars: with the default launch shapes a member equals its standalone run bit
for bit (the same kernels over the same values); with tuned shapes the
partitioned solves round differently, and members stay within 1e-12 relative
L2 of their standalone runs after 200 steps; the oracle bar is the solver's,
1e-9 relative L2.
"""
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import hit

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


def build(members, tune_cache=None, tune=False, graphs=True, **case):
    from mpi_3dcompact_amd.ensemble import Ensemble
    e = Ensemble(hit.options_text(members[0]["N"], **case), [(m["N"],) * 3 for m in members],
                 [m["mu"] for m in members])
    for i, m in enumerate(members):
        e.set_metrics(i, hit.identity_metrics(m["N"]))
        e.set_initial(i, hit.hit_initial_condition(m["N"], k0=4.0, Mt=m["Mt"], seed=m["seed"]))
    if tune:
        e.tune(tune_cache)
    e.use_graphs(graphs)
    return e


def standalone(m, steps, **case):
    from mpi_3dcompact_amd.solver import Solver
    g = Solver(hit.options_text(m["N"], mu_ref=m["mu"], **case), (m["N"],) * 3)
    g.set_metrics(hit.identity_metrics(m["N"]))
    g.set_initial(hit.hit_initial_condition(m["N"], k0=4.0, Mt=m["Mt"], seed=m["seed"]))
    g.step(steps)
    return g


SMALL = [dict(N=32, Mt=0.3, mu=1e-3, seed=1), dict(N=24, Mt=0.2, mu=5e-4, seed=2),
         dict(N=32, Mt=0.15, mu=2e-3, seed=3), dict(N=16, Mt=0.25, mu=1e-3, seed=4),
         dict(N=24, Mt=0.35, mu=8e-4, seed=5), dict(N=32, Mt=0.4, mu=4e-4, seed=6)]
C3 = dict(fd="PADE6", filt="COMPACT10", alpha_f=0.49, rk="RK4", cfl=0.5)


def test_stacked_members_equal_standalone_runs(gpu):
    _say("case: 6 members (3 x 32^3, 2 x 24^3, 16^3), RK4 PADE6 COMPACT10, 10 steps, default shapes")
    e = build(SMALL, **C3)
    e.step(10)
    info = e.info()
    assert len(info["buckets"]) == 3 and info["graph_launches"] == 9
    for i, m in enumerate(SMALL):
        g = standalone(m, 10, **C3)
        for f in CONS + ("p", "T", "mu"):
            a, b = e.field(i, f), g.field(f)
            assert np.array_equal(a, b), (i, f, np.max(np.abs(a - b)))
        t, dt, ts = e.scalars(i)
        tg, dtg, tsg = g.scalars()
        assert (t, dt, ts) == (tg, dtg, tsg), i


GRAPH_CASES = [
    dict(fd="PADE6", filt="COMPACT10", rk="RK4", cfl=0.5),
    dict(fd="PENTA10", filt="COMPACT8", rk="LSLDDRK4", cfl=0.4, filter_step=2),
    dict(fd="PADE6", filt="COMPACT8", rk="TVDRK3", ts_type="CONST_DT", dt=2e-3, cfl=None, substep_filtering=True),
    dict(fd="PENTA10", filt="COMPACT8", rk="KENRK4", cfl=0.5, les="VREMAN", les_avg="LOCAL"),
]


@pytest.mark.parametrize("case", GRAPH_CASES, ids=["rk4", "lsldd-filterstep2", "tvd-substep-constdt",
                                                   "kenrk4-vreman"])
def test_graph_replays_equal_direct_launches(gpu, case):
    members = SMALL[:4]
    _say(f"case: graphs vs direct launches, {case}")
    a = build(members, graphs=True, **case)
    b = build(members, graphs=False, **case)
    a.step(8)
    b.step(8)
    ia = a.info()
    assert ia["graph_launches"] == 7 and ia["graphs"] >= 1
    assert b.info()["graph_launches"] == 0
    for i in range(len(members)):
        for f in CONS:
            assert np.array_equal(a.field(i, f), b.field(i, f)), (i, f)
        assert a.scalars(i) == b.scalars(i)


def test_member_against_the_oracle(gpu):
    from oracle import pyoracle as P
    _say("case: 32^3 member of a stacked bucket vs the CPU oracle, 20 steps")
    e = build(SMALL[:3], **C3)
    e.step(20)
    m = SMALL[2]
    ic = hit.hit_initial_condition(m["N"], k0=4.0, Mt=m["Mt"], seed=m["seed"])
    o = P.Solver(hit.oracle_config(m["N"], mu_ref=m["mu"], **C3), hit.identity_metrics(m["N"]), ic)
    o.steps(20)
    for f in CONS:
        e_ = rel_l2(e.field(2, f), o.field(f))
        _say(f"  {f}: {e_:.2e}")
        assert e_ <= 1e-9, (f, e_)
    t, dt, ts = e.scalars(2)
    assert ts == o.timestep and abs(dt - o.dt) <= 1e-12 * o.dt


def test_tuned_ensemble_within_1e12_of_standalone_after_200_steps(gpu, tmp_path):
    """The batched-vs-standalone bar of the ensemble workload, on 12 members of
    the workload's distribution restricted to N in {32, 48, 64}."""
    members = hit.ensemble_members(12, seed=7, sizes=(32, 48, 64))
    cache = tmp_path / "shapes.txt"
    _say("case: 12 members N in {32,48,64}, tuned, 200 steps RK4 vs standalone")
    e = build(members, tune=True, tune_cache=cache, **C3)
    info = e.info()
    nb = len(info["buckets"])
    assert all(len(b["tuned"]) == 6 for b in info["buckets"])
    assert all(t["candidates"] >= 2 and not t["cached"] for b in info["buckets"] for t in b["tuned"])
    lines = [ln for ln in cache.read_text().splitlines() if " = " in ln and not ln.startswith("#")]
    assert len(lines) == 6 * nb
    e.step(200)
    worst = 0.0
    for i, m in enumerate(members):
        g = standalone(m, 200, **C3)
        for f in CONS:
            worst = max(worst, rel_l2(e.field(i, f), g.field(f)))
        assert e.scalars(i)[2] == 200
    _say(f"  worst rel L2 over members and fields: {worst:.2e}")
    assert worst <= 1e-12
    # a second ensemble takes the same shapes from the cache and reproduces the run
    e2 = build(members, tune=True, tune_cache=cache, **C3)
    i2 = e2.info()
    assert all(t["cached"] for b in i2["buckets"] for t in b["tuned"])
    text = cache.read_text().splitlines()
    assert sum(ln.startswith("#") for ln in text) == 1 and len(text) == 1 + 6 * nb
    assert [b["deriv"] + b["filter"] for b in i2["buckets"]] == [b["deriv"] + b["filter"] for b in info["buckets"]]
    e2.step(5)
    e3 = build(members, tune=True, tune_cache=cache, graphs=False, **C3)
    e3.step(5)
    for i in range(len(members)):
        assert np.array_equal(e2.field(i, "rhoE"), e3.field(i, "rhoE"))


def test_tuned_large_members_within_1e12_of_standalone_after_200_steps(gpu, tmp_path):
    """The same bar for the workload's two largest box sizes: two 96^3 members
    stacked in one bucket and one 128^3 member, M_t and MU_REF drawn from the
    workload's ranges, tuned shapes and graph replays, 200 RK4 steps."""
    r = np.random.default_rng(96128)
    members = [dict(N=n, Mt=float(r.uniform(0.1, 0.4)), mu=float(r.uniform(1.0 / 3200.0, 1.0 / 800.0)),
                    seed=int(r.integers(0, 2 ** 31 - 1))) for n in (96, 128, 96)]
    _say("case: members N = 96, 128, 96, tuned, 200 steps RK4 vs standalone")
    e = build(members, tune=True, tune_cache=tmp_path / "shapes.txt", **C3)
    assert sorted(len(b["members"]) for b in e.info()["buckets"]) == [1, 2]
    e.step(200)
    worst = 0.0
    for i, m in enumerate(members):
        g = standalone(m, 200, **C3)
        for f in CONS:
            worst = max(worst, rel_l2(e.field(i, f), g.field(f)))
        (te, dte, tse), (tg, dtg, tsg) = e.scalars(i), g.scalars()
        assert tse == tsg == 200 and abs(dte - dtg) <= 1e-12 * dtg and abs(te - tg) <= 1e-12 * tg
        g.close()
    _say(f"  worst rel L2 over members and fields: {worst:.2e}")
    assert worst <= 1e-12


def test_ensemble_refuses_what_it_does_not_batch(gpu):
    from mpi_3dcompact_amd._lib import Mc3dError
    from mpi_3dcompact_amd.ensemble import Ensemble
    with pytest.raises(Mc3dError):
        Ensemble(hit.options_text(32, les="DSM", les_avg="LOCAL"), [(32,) * 3] * 2)
    with pytest.raises(Mc3dError):
        Ensemble(hit.options_text(32, faces={"Y0": "ADIABATIC_WALL", "Y1": "ADIABATIC_WALL"}), [(32,) * 3] * 2)
    e = Ensemble(hit.options_text(32), [(32,) * 3, (16,) * 3])
    with pytest.raises(Mc3dError):
        e.step(1)  # nothing set
