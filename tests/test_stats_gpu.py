"""Running statistics on the GPU (src/Stats.cpp) against the CPU oracle, and the
files written beside a SolutionDump: VelocityStats.<step> / ThermoStats.<step>
(weight, then the fields, x fastest) and SpongeAvgDump.<step> (five sponge
averages).  Tolerance 1e-12 (FP64, 7 steps of the same flow): normwise for the
means, on the scale of the second moments for RMS (squared) and Reynolds
stresses; weights to 1e-13."""
import os
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

VEL = ("UAVG", "VAVG", "WAVG", "URMS", "VRMS", "WRMS", "UVREY", "UWREY", "VWREY")
THERMO = ("RHOAVG", "RHORMS", "PAVG", "PRMS", "TAVG", "TRMS")
CASE = dict(fd="PADE6", filt="COMPACT10", rk="KENRK4", cfl=0.5, mu_ref=1e-3)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def nrel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def gpu_solver(N, stats, ic, **case):
    from mpi_3dcompact_amd.solver import Solver
    g = Solver(hit.options_text(N, stats=stats, **case), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(ic)
    return g


@pytest.mark.parametrize("avg", ["NONE", "XI1_AVG", "XI2_AVG", "XI3_AVG"])
def test_statistics_match_the_oracle(gpu, avg):
    N = 24
    stats = dict(velocity=True, thermo=True, avg=avg, interval=2)
    ic = hit.hit_initial_condition(N, k0=2.0, Mt=0.25, seed=11)
    _say(f"case: stats {avg} interval 2, 24^3, 7 steps")
    g = gpu_solver(N, stats, ic, **CASE)
    g.step(7)
    o = P.Solver(hit.oracle_config(N, stats=stats, **CASE), hit.identity_metrics(N), ic)
    o.steps(7)
    vw, tw = g.stats_weights()
    assert vw == pytest.approx(o.stats_weight(0), rel=1e-13) and tw == pytest.approx(o.stats_weight(1), rel=1e-13)
    assert vw > 0
    # means to 1e-12 normwise; RMS = sqrt|E[x^2] - E[x]^2| and REY = E[uv] - E[u]E[v]
    # cancel over a few nearby samples, so they are compared (RMS squared) on the
    # scale of the second moments E[x^2] = AVG^2 + RMS^2
    m2 = {x: o.field(x + "AVG") ** 2 + o.field(x + "RMS") ** 2 for x in ("U", "V", "W", "RHO", "P", "T")}
    errs = {}
    for k in VEL + THERMO:
        a, b = g.field(k), o.field(k)
        if k.endswith("AVG"):
            errs[k] = nrel(a, b)
        elif k.endswith("RMS"):
            errs[k] = np.max(np.abs(a ** 2 - b ** 2)) / np.max(m2[k[:-3]])
        else:
            errs[k] = np.max(np.abs(a - b)) / np.sqrt(np.max(m2[k[0]]) * np.max(m2[k[1]]))
    _say("  worst:", max(errs.items(), key=lambda kv: kv[1]))
    for k, e in errs.items():
        assert e <= 1e-12, (k, e)


def test_statistics_files_and_their_restart(gpu, tmp_path):
    N = 16
    stats = dict(velocity=True, thermo=True, avg="NONE", interval=1)
    ic = hit.hit_initial_condition(N, k0=2.0, Mt=0.2, seed=5)
    g = gpu_solver(N, stats, ic, **CASE)
    g.step(4)
    path = tmp_path / "SolutionDump.4"
    g.write_dump(str(path))
    vw, tw = g.stats_weights()
    for stem, names, w in (("VelocityStats.4", VEL, vw), ("ThermoStats.4", THERMO, tw)):
        raw = np.fromfile(tmp_path / stem, dtype=np.float64)
        assert raw.size == 1 + len(names) * N ** 3, stem
        assert raw[0] == w
        for k, name in enumerate(names):
            assert np.array_equal(raw[1 + k * N ** 3:1 + (k + 1) * N ** 3].reshape(N, N, N), g.field(name)), name
    # a new run takes the statistics from those files
    stats2 = dict(stats, velocity_file=str(tmp_path / "VelocityStats.4"), thermo_file=str(tmp_path / "ThermoStats.4"))
    h = gpu_solver(N, stats2, ic, **CASE)
    assert h.stats_weights() == (vw, tw)
    for name in VEL + THERMO:
        assert np.array_equal(h.field(name), g.field(name)), name


def test_sponge_averages_written_beside_the_dump(gpu, tmp_path):
    from mpi_3dcompact_amd.solver import Solver
    N = 16
    faces = {"X0": "SPONGE", "X1": "SPONGE"}
    g = Solver(hit.options_text(N, faces=faces, sponge={"KIND": "RECTILINEAR", "RECTX0PERC": 0.2,
                                                        "RECTX1PERC": 0.2}, **CASE), (N, N, N))
    x, y, z = hit.box_mesh(N, faces)
    g.set_mesh(x, y, z)
    g.set_initial(hit.hit_initial_condition(N, k0=2.0, Mt=0.2, seed=2))
    g.step(2)
    g.write_dump(str(tmp_path / "SolutionDump.2"))
    raw = np.fromfile(tmp_path / "SpongeAvgDump.2", dtype=np.float64)
    assert raw.size == 5 * N ** 3
    for k, name in enumerate(("rhoAvg", "rhoUAvg", "rhoVAvg", "rhoWAvg", "rhoEAvg")):
        assert np.array_equal(raw[k * N ** 3:(k + 1) * N ** 3].reshape(N, N, N), g.field(name)), name
    assert not os.path.exists(tmp_path / "VelocityStats.2")


def test_sponge_averages_from_a_restart_file(gpu, tmp_path):
    """SPONGEAVGFROMRESTART: the sponge averages start from the five fields of a
    SpongeAvgDump file (src/SpongeBC.cpp:426-475) instead of the initial state,
    and the step relaxes towards them: one step against the oracle given the
    same averages."""
    from mpi_3dcompact_amd.solver import Solver
    N = 16
    faces = {"X0": "SPONGE", "X1": "SPONGE"}
    sponge = {"KIND": "RECTILINEAR", "RECTX0PERC": 0.25, "RECTX1PERC": 0.25}
    x, y, z = hit.box_mesh(N, faces)
    a = Solver(hit.options_text(N, faces=faces, sponge=sponge, **CASE), (N, N, N))
    a.set_mesh(x, y, z)
    a.set_initial(hit.hit_initial_condition(N, k0=2.0, Mt=0.2, seed=2))
    a.write_dump(str(tmp_path / "SolutionDump.0"))
    path = tmp_path / "SpongeAvgDump.0"
    avg = {k: a.field(k + "Avg") for k in ("rho", "rhoU", "rhoV", "rhoW", "rhoE")}
    ic = hit.hit_initial_condition(N, k0=3.0, Mt=0.3, seed=9)
    b = Solver(hit.options_text(N, faces=faces, sponge=sponge, sponge_avg_file=str(path), **CASE), (N, N, N))
    b.set_mesh(x, y, z)
    b.set_initial(ic)
    for k, v in avg.items():
        assert np.array_equal(b.field(k + "Avg"), v), k
    assert not np.array_equal(b.field("rhoU"), avg["rhoU"])  # a different state (rho is 1 in both)
    # one step relaxes towards the file's averages, as the oracle does with them
    fk = [P.FACE[k] for k in hit.face_kinds(faces)]
    sig = P.sponge_sigma(x, y, z, fk, kind=0, strength=1.0, perc=[0.25, 0.25, 0, 0, 0, 0])
    o = P.Solver(hit.oracle_config(N, faces=faces, **CASE), [b.field(n) for n in
                 ("J", "J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32", "J33")], ic)
    o.set_sponge(sig, avg)
    b.step(1)
    o.steps(1)
    for n in ("rho", "rhoU", "rhoE"):
        e = np.linalg.norm(b.field(n) - o.field(n)) / np.linalg.norm(o.field(n))
        assert e <= 1e-12, (n, e)
    # a missing file is an error at set-up
    from mpi_3dcompact_amd._lib import Mc3dError
    c = Solver(hit.options_text(N, faces=faces, sponge=sponge, sponge_avg_file=str(tmp_path / "none"), **CASE),
               (N, N, N))
    c.set_mesh(x, y, z)
    with pytest.raises(Mc3dError):
        c.set_initial(ic)
