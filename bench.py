#!/usr/bin/env python3
"""Throughput of the compact-difference LES time step on one MI355X.

Workload: decaying compressible HIT on a 256^3 periodic grid (identity
curvilinear map), FP64, PENTA10 derivatives, COMPACT8 filter (alpha_f 0.49),
KENRK4 time stepping at CFL 0.5, VREMAN SGS with LOCAL averaging,
mu_ref = 1/5000; seeded synthetic initial field (Rogallo spectrum,
M_t = 0.3).  A "step" is one full time step (5 RK stages + filter + CFL).

value = grid-point-steps per second (Mpt-steps/s), inputs resident in HBM.
roofline = the kernel with the largest share of step time, its algorithmic
bytes per launch over its mean launch duration (HIP events on the solver
stream around every launch, over as many steps again right after the timed
ones: the events cost about 2% of a step, so `value` is timed without them)
against 8 TB/s; traffic = that kernel's HBM
bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this
workload (profiles/les256_traffic.json, counters scaled by their reading for
a copy of known size), when the profiled kernel is the dominant one.
cpu_baseline = the CPU oracle (a plain C++ restatement of the reference
algorithm, OpenMP over lines) timed on this host: 1 warm-up and 3 timed steps
of the same configuration (256^3 by default) on the threads OpenMP gives it,
the thread count read back from the runtime, the host's core counts and CPU
model, and a 1-thread rate at 128^3.
effective_bandwidth = the design-independent dataflow of the reference for
this configuration (9592 B per point-step: every reference operator's distinct
full-size FP64 inputs and outputs, KenRK4 + Vreman LOCAL) times the step rate.
ensemble = a second workload of 64 HIT boxes of uneven sizes advanced together
(members of equal size stacked, one HIP graph per step, tuned line-operator
shapes): its point-step rate, the same with direct launches and with default
shapes, the members run one after another as standalone solvers, and a
parity sample of members against standalone runs.
learned_closure = the learned-vs-DSM half of the metric: "timed" runs the
DSM (LESAVERAGING LOCAL, mu 1e-4) and the learned closure (the committed
network trained on that DSM's pairs) from one 256^3 initial field here and
reports the step-time ratio; "studies" holds the errors the committed closure
studies measured over their online runs (256^3 LOCAL and XI1 averaging, 64^3),
not re-run here.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Mpt-steps/s at 256³ FP64 vs CPU; learned vs DSM LES speedup at "
          "≤2% KE error")
HBM_PEAK_GBS = 8000.0
# reference dataflow per point-step of KENRK4 + VREMAN LOCAL: stage subtotal
# 1512 B x 5 stages, KenRK4 register updates 960 B, filter + CFL 312 B, Vreman
# LOCAL 152 B x 5 stages
B_REF_PER_PT_STEP = 9592.0

CASE = dict(fd="PENTA10", filt="COMPACT8", alpha_f=0.49, rk="KENRK4", cfl=0.5,
            les="VREMAN", les_avg="LOCAL", mu_ref=1.0 / 5000.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, n):
    """HBM bytes per launch of kernel from the committed counter profile, or None."""
    path = os.path.join(ROOT, "profiles", "les256_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or n != 256:
        return None
    return t.get("traffic_bytes_per_launch")


# the learned-closure workload: DSM LES (PENTA10, COMPACT8, KENRK4, Pade test
# filter, tau_kk on, LESAVERAGING LOCAL) against the learned closure (point
# features, CLOSUREFEATURES POINT, trained on that DSM's coefficients) with its
# LOCAL-averaged DSM fallback, mu_ref 1e-4, M_t 0.3, from one initial field
CLOSURE_NET = os.path.join(ROOT, "mpi_3dcompact_amd", "networks", "hit256_local_point.net")
CLOSURE_FEATURES = "POINT"


def learned_closure_study():
    """Errors of the learned closure against the DSM from the committed closure
    studies (python -m mpi_3dcompact_amd.closure_study; training on 2e7 DSM
    pairs and the online runs take minutes, so they are not repeated here),
    each labelled with its LESAVERAGING and the steps compared."""
    out = {}
    for tag in ("256_local_point", "256_local", "256_xi1", "64_xi1"):
        path = os.path.join(ROOT, "profiles", f"closure_study_{tag}.json")
        try:
            with open(path) as f:
                d = json.load(f)
        except (OSError, ValueError):
            continue
        out[tag] = {"averaging": d.get("avg", "XI1_AVG"), "features": d.get("features", "AVERAGED"),
                    "online_steps": d["online_steps"],
                    "compared_steps": d.get("compared_steps"), "blowup_step": d.get("blowup_step"),
                    "speedup_vs_dsm": round(d["speedup_vs_dsm"], 3),
                    "learned_vs_dsm": d["learned_vs_dsm"], "constant_vs_dsm": d.get("constant_vs_dsm"),
                    "nomodel_vs_dsm": d.get("nomodel_vs_dsm"), "heldout_mse": d.get("heldout_mse"),
                    "source": f"profiles/closure_study_{tag}.json"}
    return out or None


def closure_leg(N, steps, warmup, net_path=CLOSURE_NET, features=CLOSURE_FEATURES):
    """Step time of the DSM and of the learned closure on the same N^3 initial
    field (each its own solver, `warmup` untimed steps, then `steps` timed with
    the stream synchronised around them), the resolved KE and dissipation of
    both after the timed steps, and the closure's fallback counters."""
    from mpi_3dcompact_amd import hit
    from mpi_3dcompact_amd.closure_study import case
    from mpi_3dcompact_amd.solver import Solver
    if not os.path.exists(net_path):
        raise SystemExit(f"closure network {net_path} is missing")
    ic = hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=0)
    met = hit.identity_metrics(N)
    res = {}
    for model in ("DSM", "LEARNED"):
        log(f"[bench] closure leg: {model} at {N}^3")
        g = Solver(case(N, model, 1e-4, "LOCAL", closure_file=net_path if model == "LEARNED" else None,
                        features=features if model == "LEARNED" else None), (N, N, N))
        g.set_metrics(met)
        g.set_initial(ic)
        g.step(warmup)
        g.sync()
        t0 = time.perf_counter()
        g.step(steps)
        g.sync()
        wall = time.perf_counter() - t0
        st = g.hit_stats()
        res[model] = {"ms_per_step": wall * 1e3 / steps, "ke": 0.5 * st["rho_uu"], "dissipation": st["dissipation"]}
        if model == "LEARNED":
            res[model]["closure"] = g.closure_stats()
        g.close()
    d, l = res["DSM"], res["LEARNED"]
    return {"workload": f"decaying HIT {N}^3, PENTA10 + COMPACT8 (alpha_f 0.49) + KENRK4 (CFL 0.5), mu 1e-4, "
                        "M_t 0.3; DSM (Pade test filter, tau_kk, LESAVERAGING LOCAL) vs the learned closure "
                        f"(CLOSUREFEATURES {features}, LOCAL-averaged DSM fallback on), network "
                        + os.path.relpath(net_path, ROOT),
            "steps": steps, "warmup": warmup,
            "ms_per_step": {"DSM": round(d["ms_per_step"], 3), "LEARNED": round(l["ms_per_step"], 3)},
            "speedup_vs_dsm": round(d["ms_per_step"] / l["ms_per_step"], 3),
            "ke_rel_diff": abs(l["ke"] / d["ke"] - 1.0),
            "dissipation_rel_diff": abs(l["dissipation"] / d["dissipation"] - 1.0),
            "fallback": l["closure"]}


def host_cpu():
    """Cores this process may use and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline(n_cpu, steps=3, n_one=128):
    """Oracle step rate on the host (test infrastructure, used only as the
    checker/baseline): 1 warm-up and `steps` timed steps at n_cpu^3 on the
    threads OpenMP gives the oracle, then one step at n_one^3 on one thread."""
    from mpi_3dcompact_amd import hit
    from oracle import pyoracle as P
    threads = P.max_threads()
    ic = hit.hit_initial_condition(n_cpu, k0=4.0, Mt=0.3, seed=0)
    s = P.Solver(hit.oracle_config(n_cpu, **CASE), hit.identity_metrics(n_cpu), ic)
    del ic
    t0 = time.perf_counter()
    s.steps(1)
    warm = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.steps(steps)
    dt = (time.perf_counter() - t0) / steps
    del s
    log(f"[bench] CPU oracle {n_cpu}^3 on {threads} threads: warm-up {warm:.1f} s, {dt:.2f} s per step")
    one = None
    if n_one:
        ic = hit.hit_initial_condition(n_one, k0=4.0, Mt=0.3, seed=0)
        s1 = P.Solver(hit.oracle_config(n_one, **CASE), hit.identity_metrics(n_one), ic)
        P.set_threads(1)
        try:
            s1.steps(1)  # one step alone: the whole step is the sample
            t0 = time.perf_counter()
            s1.steps(1)
            dt1 = time.perf_counter() - t0
        finally:
            P.set_threads(threads)
        one = {"value": n_one ** 3 / dt1 / 1e6, "unit": "Mpt-steps/s", "cores": 1,
               "sample": f"1 step at {n_one}^3 after 1 warm-up step ({dt1:.1f} s)"}
        log(f"[bench] CPU oracle {n_one}^3 on 1 thread: {dt1:.2f} s per step")
    return {"value": n_cpu ** 3 / dt / 1e6, "unit": "Mpt-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} KENRK4 steps after 1 warm-up step of the same PENTA10/COMPACT8/VREMAN LOCAL "
                      f"case at {n_cpu}^3 ({dt:.2f} s per step), OpenMP over independent lines on "
                      f"{threads} threads (omp_get_max_threads)",
            "host": host_cpu(), "one_core": one}


# ensemble workload: 64 decaying-HIT boxes, N drawn from {32, 48, 64, 96, 128},
# M_t ~ U[0.1, 0.4], MU_REF ~ U[1/3200, 1/800] (hit.ensemble_members, seed 2024),
# PADE6 + COMPACT10 (alpha_f 0.49) + RK4 at CFL 0.5, each box its own dt
ENS_CASE = dict(fd="PADE6", filt="COMPACT10", alpha_f=0.49, rk="RK4", cfl=0.5)
# reference dataflow per point-step of plain RK4: stage subtotal 1512 B x 4,
# RK4 register updates 680 B, filter + CFL 312 B
B_REF_RK4 = 7040.0


def ensemble_leg(steps, warmup, check_members=3, cache=None, step_only=False):
    """Ensemble throughput (graph replays, tuned shapes) with the same ensemble
    without graphs and without tuning, the members run one after another as
    standalone solvers, and a parity sample of members against standalone runs.
    cache: tuning cache file; step_only: the tuned graph-replayed steps alone."""
    import numpy as np
    from mpi_3dcompact_amd import hit
    from mpi_3dcompact_amd.ensemble import Ensemble
    from mpi_3dcompact_amd.solver import Solver
    members = hit.ensemble_members(64, seed=2024)
    pts = sum(m["N"] ** 3 for m in members)
    log(f"[bench] ensemble: {len(members)} members, {pts / 1e6:.1f} M points; generating initial fields")
    ics = [hit.hit_initial_condition(m["N"], k0=4.0, Mt=m["Mt"], seed=m["seed"]) for m in members]
    opts = hit.options_text(members[0]["N"], **ENS_CASE)

    def make(tune):
        e = Ensemble(opts, [(m["N"],) * 3 for m in members], [m["mu"] for m in members])
        for i, m in enumerate(members):
            e.set_metrics(i, hit.identity_metrics(m["N"]))
            e.set_initial(i, ics[i])
        t0 = time.perf_counter()
        if tune:
            e.tune(cache)
        return e, time.perf_counter() - t0

    def timed(e, k):
        e.sync()
        t0 = time.perf_counter()
        e.step(k)
        e.sync()
        return time.perf_counter() - t0

    res = {}
    if step_only:
        e, tune_s = make(True)
        e.step(warmup)
        wall = timed(e, steps)
        e.close()
        value = pts * steps / wall / 1e6
        return {"points": pts, "value": round(value, 3), "unit": "Mpt-steps/s",
                "ms_per_step": round(wall * 1e3 / steps, 3), "steps": steps, "warmup": warmup,
                "tune_seconds": round(tune_s, 2)}
    # default shapes, graphs
    e, _ = make(False)
    e.step(warmup)
    res["untuned_graph_s"] = timed(e, steps)
    e.close()
    # tuned shapes: graphs, then direct launches on the same ensemble
    e, tune_s = make(True)
    info = e.info()
    log(f"[bench] ensemble tuned in {tune_s:.1f} s")
    e.step(warmup)
    wall = timed(e, steps)
    e.use_graphs(False)
    res["tuned_direct_s"] = timed(e, steps)
    total_steps = warmup + 2 * steps
    # parity sample: the largest, smallest and a middle size against standalone runs
    sizes = sorted({m["N"] for m in members})
    pick = [next(i for i, m in enumerate(members) if m["N"] == n)
            for n in (sizes[0], sizes[len(sizes) // 2], sizes[-1])][:check_members]
    worst = 0.0
    for i in pick:
        m = members[i]
        g = Solver(hit.options_text(m["N"], mu_ref=m["mu"], **ENS_CASE), (m["N"],) * 3)
        g.set_metrics(hit.identity_metrics(m["N"]))
        g.set_initial(ics[i])
        g.step(total_steps)
        for f in ("rho", "rhoU", "rhoV", "rhoW", "rhoE"):
            a, b = e.field(i, f), g.field(f)
            worst = max(worst, float(np.linalg.norm(a - b) / np.linalg.norm(b)))
        g.close()
    finite = all(np.isfinite(e.field(i, "rho")).all() for i in range(len(members)))
    e.close()
    if not finite:
        raise SystemExit("non-finite density in the ensemble")
    # the members one after another, each as a standalone solver
    log("[bench] ensemble members as standalone solvers")
    alone = 0.0
    per = 3
    for i, m in enumerate(members):
        g = Solver(hit.options_text(m["N"], mu_ref=m["mu"], **ENS_CASE), (m["N"],) * 3)
        g.set_metrics(hit.identity_metrics(m["N"]))
        g.set_initial(ics[i])
        g.step(2)
        g.sync()
        t0 = time.perf_counter()
        g.step(per)
        g.sync()
        alone += (time.perf_counter() - t0) / per
        g.close()
    value = pts * steps / wall / 1e6
    return {
        "workload": "64 decaying-HIT boxes, N from {32,48,64,96,128}, PADE6 + COMPACT10 (alpha_f 0.49) + RK4 "
                    "(CFL 0.5, own dt per box), M_t U[0.1,0.4], MU_REF U[1/3200,1/800], seed 2024",
        "points": pts, "buckets": len(info["buckets"]),
        "value": round(value, 3), "unit": "Mpt-steps/s", "ms_per_step": round(wall * 1e3 / steps, 3),
        "steps": steps, "warmup": warmup,
        "direct_launch_value": round(pts * steps / res["tuned_direct_s"] / 1e6, 3),
        "untuned_value": round(pts * steps / res["untuned_graph_s"] / 1e6, 3),
        "standalone_sequential_value": round(pts / alone / 1e6, 3),
        "batched_over_standalone": round((pts * steps / wall) / (pts / alone), 3),
        "tune_seconds": round(tune_s, 2),
        "effective_bandwidth": {"bytes_per_pt_step": B_REF_RK4,
                                "achieved": round(B_REF_RK4 * value * 1e6 / 1e9, 1), "unit": "GB/s",
                                "frac": round(B_REF_RK4 * value * 1e6 / 1e9 / HBM_PEAK_GBS, 4)},
        "parity_vs_standalone": {"members": pick, "steps": total_steps, "max_rel_l2": worst},
        "shapes": [{"dims": b["dims"][0], "members": len(b["members"]), "deriv": b["deriv"], "filter": b["filter"]}
                   for b in info["buckets"]],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="grid points per direction")
    ap.add_argument("--walls", action="store_true",
                    help="adiabatic walls on every face (Dirichlet lines in every direction) instead of the periodic box")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU oracle steps after one warm-up")
    ap.add_argument("--cpu-one-core-n", type=int, default=128, help="grid of the 1-thread CPU sample (0: skip)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble workload")
    ap.add_argument("--no-closure", action="store_true", help="skip the learned-closure timing leg")
    ap.add_argument("--closure-only", action="store_true", help="run the learned-closure timing leg alone")
    ap.add_argument("--closure-steps", type=int, default=5)
    ap.add_argument("--closure-net", default=CLOSURE_NET, help="network file of the learned-closure leg")
    ap.add_argument("--closure-features", default=CLOSURE_FEATURES, choices=("AVERAGED", "POINT"),
                    help="CLOSUREFEATURES the network was trained with")
    ap.add_argument("--ensemble-only", action="store_true", help="run the ensemble workload alone")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-cache", default=None, help="line-operator tuning cache file")
    ap.add_argument("--ensemble-step-only", action="store_true",
                    help="with --ensemble-only: the tuned, graph-replayed steps alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from mpi_3dcompact_amd import hit
    from mpi_3dcompact_amd.solver import Solver

    if args.closure_only:
        print(json.dumps({"learned_closure": closure_leg(args.n, args.closure_steps, 2, args.closure_net,
                                                         args.closure_features)}), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_leg(args.ensemble_steps, 4, cache=args.ensemble_cache,
                                                   step_only=args.ensemble_step_only)}), flush=True)
        return

    N = args.n
    log(f"[bench] generating {N}^3 HIT initial field")
    ic = hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=0)
    met = hit.identity_metrics(N)
    faces = {f: "ADIABATIC_WALL" for f in hit.FACE_NAMES} if args.walls else None
    g = Solver(hit.options_text(N, faces=faces, **CASE), (N, N, N))
    g.set_metrics(met)
    g.set_initial(ic)
    del met
    log(f"[bench] warmup {args.warmup} steps")
    g.step(args.warmup)
    g.sync()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    wall = time.perf_counter() - t0
    # per-kernel times from HIP events around every launch, over as many
    # further steps (the events add about 2% to a step, so they stay out of
    # the steps that give `value`)
    log(f"[bench] {args.steps} more steps with per-launch HIP events")
    g.timing(True)
    t1 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    wall_events = time.perf_counter() - t1
    rep = g.timing_report()
    g.timing(False)
    ms_per_step = wall * 1e3 / args.steps
    value = N ** 3 * args.steps / wall / 1e6
    kern_ms = sum(r["ms"] for r in rep)
    dom = max(rep, key=lambda r: r["ms"])
    avg_ms = dom["ms"] / dom["count"]
    bytes_per_launch = dom["bytes"] / dom["count"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    for r in sorted(rep, key=lambda r: -r["ms"]):
        gbs = r["bytes"] / (r["ms"] * 1e-3) / 1e9 if r["ms"] > 0 else 0.0
        log(f"[bench] {r['kernel']:28s} launches {r['count']:5d}  {r['ms'] / args.steps:8.2f} ms/step "
            f"{100 * r['ms'] / kern_ms:5.1f}%  {gbs:7.0f} GB/s")
    t_final, dt_final, ts = g.scalars()
    rho = g.field("rho")
    if not np.isfinite(rho).all():
        raise SystemExit("non-finite density after the timed steps")
    g.close()

    cpu = None
    if not args.no_cpu and not args.walls:  # (the CPU sample is the periodic box)
        log(f"[bench] CPU oracle baseline at {args.cpu_n}^3")
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps, args.cpu_one_core_n)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"full plain LES, decaying HIT {N}^3 "
                               f"{'between adiabatic walls on every face' if args.walls else 'periodic'}, PENTA10 + COMPACT8 "
                               f"(alpha_f 0.49) + KENRK4 (CFL 0.5) + VREMAN LOCAL, mu 1/5000",
                   "grid": [N, N, N], "points": N ** 3, "rk_stages_per_step": 5},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if args.walls else measured_traffic(dom["kernel"], N), "bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": round(avg_ms, 4),
                     "share_of_step": round(dom["ms"] / kern_ms, 3),
                     "events": f"HIP events around every launch over {args.steps} further steps "
                               f"({wall_events * 1e3 / args.steps:.3f} ms/step with the events, "
                               f"{kern_ms / args.steps:.3f} ms/step of kernels)"},
        "cpu_baseline": cpu,
        "effective_bandwidth": {"bytes_per_pt_step": B_REF_PER_PT_STEP,
                                "achieved": round(B_REF_PER_PT_STEP * value * 1e6 / 1e9, 1),
                                "unit": "GB/s",
                                "frac": round(B_REF_PER_PT_STEP * value * 1e6 / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    lc = {} if args.no_closure else {"timed": closure_leg(N, args.closure_steps, 2, args.closure_net,
                                                          args.closure_features)}
    lc["studies"] = learned_closure_study()
    out["learned_closure"] = lc
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ensemble_steps, 4, cache=args.ensemble_cache)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
