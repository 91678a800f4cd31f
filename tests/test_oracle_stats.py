"""Pins of the oracle's running statistics (src/Stats.cpp:121-345) by what the
mathematics fixes.  The recursive update is a time-weighted average: after
updates at times t_1 < t_2 < ... with weights w_i = t_i - t_{i-1} (t_0 the
initial time; updates with w = 0 are skipped),
  AVG = sum w x / sum w,  RMS = sqrt|sum w x^2 / sum w - AVG^2|,
  REY = sum w u v / sum w - AVG_u AVG_v,
recomputed here from the sampled fields (U V W, the unfiltered register rho2,
p, T).  An XI1 / XI3 average replaces every running moment by its line mean
after each update, so the results are the line means of the moments; the
reference's XI2 branch averages a scratch array, so XI2 equals NONE bit for bit.
"""
import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

VEL = ("UAVG", "VAVG", "WAVG", "URMS", "VRMS", "WRMS", "UVREY", "UWREY", "VWREY")
THERMO = ("RHOAVG", "RHORMS", "PAVG", "PRMS", "TAVG", "TRMS")
CASE = dict(fd="PADE6", filt="COMPACT10", rk="TVDRK3", cfl=0.5, mu_ref=2e-3)


def run(N, steps, avg="NONE", interval=1, record=False, seed=3):
    ic = hit.hit_initial_condition(N, k0=2.0, Mt=0.2, seed=seed)
    stats = dict(velocity=True, thermo=True, avg=avg, interval=interval)
    o = P.Solver(hit.oracle_config(N, stats=stats, **CASE), hit.identity_metrics(N), ic)
    samples = []
    for s in range(steps):
        o.steps(1)
        if record and o.timestep % interval == 0:
            samples.append((o.time, {k: o.field(k) for k in ("U", "V", "W", "rho2", "p", "T")}))
    return o, samples


def moments(samples, t0, reduce=lambda a: a):
    w_sum, acc = 0.0, {}
    t_prev = t0
    for t, f in samples:
        w = t - t_prev
        t_prev = t
        if w == 0.0:
            continue
        w_sum += w
        terms = {"U": f["U"], "V": f["V"], "W": f["W"], "R": f["rho2"], "P": f["p"], "T": f["T"],
                 "UU": f["U"] ** 2, "VV": f["V"] ** 2, "WW": f["W"] ** 2, "RR": f["rho2"] ** 2,
                 "PP": f["p"] ** 2, "TT": f["T"] ** 2, "UV": f["U"] * f["V"], "UW": f["U"] * f["W"],
                 "VW": f["V"] * f["W"]}
        for k, v in terms.items():
            acc[k] = acc.get(k, 0.0) + w * v
    m = {k: reduce(v / w_sum) for k, v in acc.items()}
    out = {"UAVG": m["U"], "VAVG": m["V"], "WAVG": m["W"], "RHOAVG": m["R"], "PAVG": m["P"], "TAVG": m["T"]}
    for a, b in (("U", "URMS"), ("V", "VRMS"), ("W", "WRMS"), ("R", "RHORMS"), ("P", "PRMS"), ("T", "TRMS")):
        out[b + "2"] = m[a + a] - m[a] ** 2          # compared as squares (cancellation)
        out[b + "_scale"] = m[a + a]
    out["UVREY"] = m["UV"] - m["U"] * m["V"]
    out["UWREY"] = m["UW"] - m["U"] * m["W"]
    out["VWREY"] = m["VW"] - m["V"] * m["W"]
    return out, w_sum


def check(o, want, w_sum):
    assert o.stats_weight(0) == pytest.approx(w_sum, rel=1e-13)
    assert o.stats_weight(1) == pytest.approx(w_sum, rel=1e-13)
    for k in VEL + THERMO:
        g = o.field(k)
        if k.endswith("RMS"):
            # sqrt|E[x^2] - E[x]^2| loses digits to cancellation; compare squares on E[x^2]'s scale
            np.testing.assert_allclose(g ** 2, np.abs(want[k + "2"]),
                                       atol=1e-11 * np.max(np.abs(want[k + "_scale"])), err_msg=k)
        else:
            scale = np.max(np.abs(want[k])) + (1e-300 if k.endswith("AVG") else np.max(np.abs(want["UAVG"])) ** 2)
            np.testing.assert_allclose(g, want[k], rtol=0, atol=1e-12 * scale, err_msg=k)


@pytest.mark.parametrize("interval", [1, 3])
def test_statistics_are_time_weighted_moments(interval):
    o, samples = run(16, 7, interval=interval, record=True)
    assert len(samples) == 7 // interval
    want, w_sum = moments(samples, 0.0)
    check(o, want, w_sum)


@pytest.mark.parametrize("avg,axis", [("XI1_AVG", 2), ("XI3_AVG", 0)])
def test_line_averaged_statistics_are_line_means_of_the_moments(avg, axis):
    o, samples = run(16, 5, avg=avg, record=True)
    want, w_sum = moments(samples, 0.0, reduce=lambda a: np.broadcast_to(a.mean(axis=axis, keepdims=True), a.shape))
    check(o, want, w_sum)
    # and the fields are constant along the line
    for k in VEL + THERMO:
        f = o.field(k)
        assert np.allclose(f, f.mean(axis=axis, keepdims=True), rtol=0, atol=1e-14 * max(1.0, np.max(np.abs(f)))), k


def test_xi2_average_leaves_the_fields_as_none_does():
    a, _ = run(16, 4, avg="XI2_AVG")
    b, _ = run(16, 4, avg="NONE")
    for k in VEL + THERMO:
        assert np.array_equal(a.field(k), b.field(k)), k


def test_first_step_adds_no_weight():
    # the first step keeps the initial time (calcDtFromCFL), so its update has dt = 0 and is skipped
    o, samples = run(16, 1, record=True)
    assert o.time == 0.0 and o.stats_weight(0) == 0.0
    assert np.all(o.field("UAVG") == 0.0)
