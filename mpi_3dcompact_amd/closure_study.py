"""Learned-closure study: DSM data -> GPU training -> online LES against the DSM.

    python -m mpi_3dcompact_amd.closure_study [--N 256] [--online-steps 1000] ...

1. A DSM LES (PENTA10, COMPACT8, KENRK4, DSMTESTFILTER 1, LESAVERAGING --avg,
   DSMUSETAUKK, DSMCOEFFRANGE) of decaying HIT from a training seed; at evenly
   spaced snapshots from step 0 to --gen-steps the DSM's coefficients
   (C, C / Pr_t) and the closure features of strided points are sampled on the
   GPU, the features averaged as the DSM averages its contractions, or taken
   at each point (--features POINT).
2. Standardisation from the samples (targets winsorised at the 0.1 / 99.9
   percentiles), the trusted input range = the sampled range widened by 10%
   of its span, then Adam training with the library's BF16-MFMA kernels.
3. From a different seed, the same LES is run with the DSM, with the learned
   closure (DSM fallback on), with constant coefficients (the network with
   zero weights: the training-mean C and C / Pr_t) and without a model; the
   resolved kinetic energy <rho u.u> / 2 and dissipation <2 mu S*:grad u> /
   <rho> (the turbdata.out statistics) are recorded every --stats-every steps,
   and the step time of each run is measured with the run's own stream
   synchronised around the timed steps.
Prints one JSON object (also written to --out).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import hit
from .closure import Closure
from .solver import Solver


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def case(N, model, mu_ref, avg="LOCAL", closure_file=None, features=None):
    return hit.options_text(N, fd="PENTA10", filt="COMPACT8", alpha_f=0.49, rk="KENRK4", cfl=0.5, mu_ref=mu_ref,
                            les=model, les_avg=avg, dsm_filter=1, use_taukk=True,
                            coeff_range=(model == "DSM"), closure_file=closure_file, closure_features=features)


def make_solver(N, model, mu_ref, seed, k0, Mt, avg="LOCAL", features=None):
    g = Solver(case(N, model, mu_ref, avg, features=features), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(hit.hit_initial_condition(N, k0=k0, Mt=Mt, seed=seed))
    return g


def generate(args, torch):
    N = args.N
    npts = N ** 3
    g = make_solver(N, "DSM", args.mu, args.train_seed, args.k0, args.Mt, args.avg, args.features)
    # snapshots at steps 0, every, ..., gen_steps: the pairs span the decay the
    # closure is used over (the online runs start from another realisation)
    nsnap = args.snapshots + 1
    per = int(args.samples // nsnap)
    stride = max(1, npts // per)
    per = min(per, (npts - 1) // stride + 1)
    X = torch.empty((per * nsnap, 7), dtype=torch.float32, device="cuda")
    T = torch.empty((per * nsnap, 2), dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(args.train_seed)
    every = max(1, args.gen_steps // args.snapshots)
    t0 = time.time()
    for s in range(nsnap):
        if s:
            g.step(every)
        off = int(rng.integers(0, npts - (per - 1) * stride))
        st = g.hit_stats()
        if not all(np.isfinite(v) for v in st.values()):
            raise RuntimeError(f"the DSM run is not finite at step {s * every}")
        g.sample_closure(X[s * per:], T[s * per:], per, off, stride)
        _say(f"[closure] DSM step {s * every}: sampled {per} points (stride {stride}), "
             f"KE {0.5 * st['rho_uu']:.5e}, {time.time() - t0:.1f} s")
    del g
    return X, T


def standardise(X, T, torch):
    ok = torch.isfinite(X).all(dim=1) & torch.isfinite(T).all(dim=1)
    X, T = X[ok].contiguous(), T[ok].contiguous()
    if X.shape[0] < 100:
        raise RuntimeError(f"only {X.shape[0]} finite training pairs")
    lo_t = torch.quantile(T[:: max(1, T.shape[0] // 1000000)].double(), 0.001, dim=0).float()
    hi_t = torch.quantile(T[:: max(1, T.shape[0] // 1000000)].double(), 0.999, dim=0).float()
    T = torch.minimum(torch.maximum(T, lo_t), hi_t)
    xm, xs = X.double().mean(0), X.double().std(0)
    tm, ts = T.double().mean(0), T.double().std(0)
    # a feature without spread (log Delta on a uniform grid) is passed through
    # unscaled; its trusted range is then the single value it took
    xs = torch.where(xs > 0, xs, torch.ones_like(xs))
    ts = torch.where(ts > 0, ts, torch.ones_like(ts))
    xmin, xmax = X.min(0).values.double(), X.max(0).values.double()
    span = xmax - xmin
    sc = dict(in_mean=xm.cpu().numpy(), in_std=xs.cpu().numpy(), out_mean=tm.cpu().numpy(),
              out_std=ts.cpu().numpy(), lo=(xmin - 0.1 * span).cpu().numpy(), hi=(xmax + 0.1 * span).cpu().numpy())
    Xn = ((X - xm.float()) / xs.float()).contiguous()
    Tn = ((T - tm.float()) / ts.float()).contiguous()
    return Xn, Tn, sc, int(ok.sum().item())


def online(args, model, net, torch):
    N = args.N
    g = Solver(case(N, model, args.mu, args.avg, features=args.features), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(hit.hit_initial_condition(N, k0=args.k0, Mt=args.Mt, seed=args.eval_seed))
    if net is not None:
        g.set_closure(net)
    hist, steps, t_run, blowup = [], 0, 0.0, None
    st = g.hit_stats()
    hist.append((0, st["time"], 0.5 * st["rho_uu"], st["dissipation"], 0.0))
    while steps < args.online_steps:
        n = min(args.stats_every, args.online_steps - steps)
        g.sync()
        t0 = time.perf_counter()
        g.step(n)
        g.sync()
        t_run += time.perf_counter() - t0
        steps += n
        st = g.hit_stats()
        if not all(np.isfinite(v) for v in st.values()):
            if not args.allow_blowup:
                raise RuntimeError(f"{model}: non-finite statistics at step {steps}")
            _say(f"[closure] {model}: non-finite statistics at step {steps}; the run ends there")
            blowup = steps
            break
        # (step, time, KE, dissipation, wall ms per step so far)
        hist.append((steps, st["time"], 0.5 * st["rho_uu"], st["dissipation"], 1e3 * t_run / steps))
        if steps % (10 * args.stats_every) == 0 or steps == args.online_steps:
            _say(f"[closure] {model}{'' if net is None else '(' + net_label(net) + ')'} step {steps}: "
                 f"KE {hist[-1][2]:.6e} eps {hist[-1][3]:.6e} {1e3 * t_run / steps:.2f} ms/step")
    extra = g.closure_stats() if model == "LEARNED" else None
    del g
    return dict(hist=hist, ms_per_step=1e3 * t_run / steps, closure=extra, blowup_step=blowup)


def net_label(net):
    return getattr(net, "label", "net")


def errors(run, ref):
    """Largest relative KE and dissipation differences over the steps both runs
    reached with finite statistics."""
    n = min(len(run["hist"]), len(ref["hist"]))
    h, r = np.array(run["hist"][:n]), np.array(ref["hist"][:n])
    ke = np.max(np.abs(h[:, 2] / r[:, 2] - 1.0))
    eps = np.max(np.abs(h[:, 3] / r[:, 3] - 1.0))
    return float(ke), float(eps)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--mu", type=float, default=1e-4)
    ap.add_argument("--Mt", type=float, default=0.3)
    ap.add_argument("--k0", type=float, default=4.0)
    ap.add_argument("--avg", default="LOCAL", help="LESAVERAGING of the DSM (and of its fallback); the "
                    "LOCAL-averaged DSM diverges after a few hundred steps at 256^3 (XI1_AVG runs 1000)")
    ap.add_argument("--features", choices=("AVERAGED", "POINT"), default=None,
                    help="CLOSUREFEATURES: the network's inputs averaged as LESAVERAGING averages the DSM's "
                         "contractions (default), or taken at each point")
    ap.add_argument("--train-seed", type=int, default=1)
    ap.add_argument("--eval-seed", type=int, default=0)
    ap.add_argument("--gen-steps", type=int, default=1000, help="DSM steps of the training run (sampled at 0, "
                    "gen_steps / snapshots, ..., gen_steps)")
    ap.add_argument("--snapshots", type=int, default=8)
    ap.add_argument("--samples", type=float, default=2e7)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--online-steps", type=int, default=1000)
    ap.add_argument("--stats-every", type=int, default=25)
    ap.add_argument("--skip-baselines", action="store_true", help="online DSM and learned runs only")
    ap.add_argument("--out", default=None)
    ap.add_argument("--allow-blowup", action="store_true", help="end an online run at its first non-finite "
                    "statistics instead of failing; errors are then taken over the steps both runs reached")
    ap.add_argument("--save-net", default=None, help="write the trained network (mc3d_closure_save) here")
    args = ap.parse_args(argv)
    import torch

    t0 = time.time()
    X, T = generate(args, torch)
    Xn, Tn, sc, n = standardise(X, T, torch)
    del X, T
    # held-out 2%: the constant predictor has standardised MSE 1 per output
    n_hold = max(1, n // 50)
    perm = torch.randperm(n, generator=torch.Generator().manual_seed(args.train_seed)).cuda()
    hold, train = perm[:n_hold], perm[n_hold:]
    Xh, Th = Xn[hold].contiguous(), Tn[hold].contiguous()
    Xt, Tt = Xn[train].contiguous(), Tn[train].contiguous()
    del Xn, Tn
    net = Closure(args.train_seed)
    net.set_scaling(sc["in_mean"], sc["in_std"], sc["out_mean"], sc["out_std"], sc["lo"], sc["hi"])
    t_tr = time.time()
    losses = net.train(Xt, Tt, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.train_seed)
    torch.cuda.synchronize()
    t_tr = time.time() - t_tr
    Yh = net.infer(Xh)
    mse = ((Yh - Th) ** 2).mean(0).double().cpu().numpy()
    _say(f"[closure] trained on {Xt.shape[0]} pairs in {t_tr:.1f} s; epoch losses "
         + " ".join(f"{v:.4f}" for v in losses) + f"; held-out MSE {mse} (constant predictor: 1)")
    del Xt, Tt, Xh, Th
    net.label = "learned"
    if args.save_net:
        net.save(args.save_net)
        _say(f"[closure] network written to {args.save_net}")
    const = Closure(0)
    const.set_params(np.zeros(8962, np.float32))
    const.set_scaling(sc["in_mean"], sc["in_std"], sc["out_mean"], sc["out_std"], np.full(7, -1e30),
                      np.full(7, 1e30))
    const.label = "constant"
    runs = {"DSM": online(args, "DSM", None, torch)}
    runs["LEARNED"] = online(args, "LEARNED", net, torch)
    if not args.skip_baselines:
        runs["CONSTANT"] = online(args, "LEARNED", const, torch)
        runs["NOMODEL"] = online(args, "NOMODEL", None, torch)
    res = dict(N=args.N, avg=args.avg, features=args.features or "AVERAGED", mu=args.mu, Mt=args.Mt, k0=args.k0, gen_steps=args.gen_steps,
               train_seed=args.train_seed, eval_seed=args.eval_seed, samples=n, epochs=args.epochs, batch=args.batch, lr=args.lr, train_seconds=t_tr,
               epoch_losses=losses, heldout_mse=mse.tolist(), scaling={k: v.tolist() for k, v in sc.items()},
               online_steps=args.online_steps, ms_per_step={k: v["ms_per_step"] for k, v in runs.items()},
               fallback=runs["LEARNED"]["closure"], total_seconds=time.time() - t0)
    for k, v in runs.items():
        if k != "DSM":
            ke, eps = errors(v, runs["DSM"])
            res[f"{k.lower()}_vs_dsm"] = dict(ke_max_rel=ke, dissipation_max_rel=eps)
    res["speedup_vs_dsm"] = runs["DSM"]["ms_per_step"] / runs["LEARNED"]["ms_per_step"]
    # over the steps the DSM run reached (its LOCAL-averaged form can diverge first)
    last = runs["DSM"]["hist"][-1][0]
    lw = [h for h in runs["LEARNED"]["hist"] if h[0] <= last]
    dms, lms = runs["DSM"]["hist"][-1][4], lw[-1][4]
    res["compared_window"] = dict(steps=last, dsm_ms_per_step=dms, learned_ms_per_step=lms,
                                  speedup_vs_dsm=dms / lms if lms > 0 else None)
    res["blowup_step"] = {k: v["blowup_step"] for k, v in runs.items()}
    res["compared_steps"] = {k: v["hist"][min(len(v["hist"]), len(runs["DSM"]["hist"])) - 1][0]
                             for k, v in runs.items() if k != "DSM"}
    res["history"] = {k: v["hist"] for k, v in runs.items()}
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
