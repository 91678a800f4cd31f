"""Long decaying-HIT runs on one GPU with the statistics the reference's HIT
driver writes to turbdata.out (tests/hit/Samtaney_D4/MPI_3DCompact.cpp:372-557)
and velocity energy spectra.

    python -m mpi_3dcompact_amd.hit_study --case les256 --out les256_500steps.json
    python -m mpi_3dcompact_amd.hit_study --case samtaney_d4 --out samtaney_d4_128.json

Cases:
  les256       256^3 PENTA10 + COMPACT8 (alpha_f 0.49) + KENRK4 (CFL 0.5) +
               VREMAN (LESAVERAGING LOCAL), mu 1/5000, 500 steps; spectra at
               steps 0, 250 and 500.
  samtaney_d4  the reference's Samtaney D4 case (tests/hit/Samtaney_D4/
               MPI_3DCompact.in: 128^3, PADE6, COMPACT10 alpha_f 0.495,
               LSLDDRK4, CFL 0.5, MU_REF 3.3e-4, no SGS model; its driver
               starts from rho = 1, p = 11.11111111/gamma and velocity files of
               M_t 0.3 peaked at k0 = 8, MPI_3DCompact.cpp:191-231) run to
               t = 7.2 tau_0 from this package's seeded field with the same
               k0, M_t and thermodynamic state, with K/K_0, M_t/M_t0 and
               rho'/M_t0^2 at t/tau_0 = 1, 2, 4, 7 beside the reference's
               published rows.  The velocity files themselves are not in the
               reference, so the random phases differ: this compares the
               decay, not values.

Statistics rows: step, then the 11 turbdata.out columns (t, tau, Re_lambda,
M_t, eps, u', eta_K, rho', dil', omega', <rho u.u>), from mc3d_solver_hit_stats.
E(k) = sum over |k'| in [k - 1/2, k + 1/2) of |u_hat(k')|^2 / 2 with
u_hat = FFT(u) / N^3, so that sum_k E(k) = <u.u> / 2.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import hit
from .solver import Solver

COLS = ("time", "tau", "re_lambda", "mach_t", "dissipation", "uprime", "kolmogorov", "rhoprime", "dilprime",
        "vortprime", "rho_uu")

# rows 1, 13, 25, 47, 78 of tests/hit/Samtaney_D4/turbdata.out (columns as COLS)
SAMTANEY_D4_REF = [
    [0.040022993696092267, 0.54816456263397462, 68.195906284722525, 0.29697333778519602, 0.16508627460854092,
     0.57159575004868302, 0.015156663614987066, 0.024960990516379698, 0.83924827735201335, 8.8902790553005016,
     0.97931307179335669],
    [0.55952688452168431, 0.63864731781435102, 38.282799150642944, 0.25407228809916182, 0.28188052518161388,
     0.49061277733253827, 0.013308315974537063, 0.029927721945887605, 0.69443134406621276, 11.636375073123517,
     0.71827701250051312],
    [1.0981828130065294, 0.8186395405208996, 27.865907623126528, 0.19750295068126633, 0.19590003199890496,
     0.382742976389223, 0.014635200981265925, 0.023038244247946638, 0.55490163262364001, 9.6777559566439812,
     0.43808254263740959],
    [2.1783288128072771, 1.2951523811030019, 19.838527389019898, 0.12442753839955309, 0.061715793526478879,
     0.24192406924506624, 0.019608136809292843, 0.016649133357330596, 0.40158190221323037, 5.4127307691268598,
     0.17535680871346801],
    [3.8136926527260333, 2.0540229680569726, 15.664127166846907, 0.078353713842907957, 0.015392062863029243,
     0.15254383188580986, 0.027788383408498225, 0.012933063064719802, 0.28720500242339764, 2.6905905438394599,
     0.069763623619035708],
]

CASES = {
    "les256": dict(N=256, opts=dict(fd="PENTA10", filt="COMPACT8", alpha_f=0.49, rk="KENRK4", cfl=0.5,
                                    les="VREMAN", les_avg="LOCAL", mu_ref=1.0 / 5000.0),
                   ic=dict(k0=4.0, Mt=0.3, c0=1.0), steps=500, spectra=(0, 250, 500), every=10),
    "samtaney_d4": dict(N=128, opts=dict(fd="PADE6", filt="COMPACT10", alpha_f=0.495, rk="LSLDDRK4", cfl=0.5,
                                         les="NOMODEL", mu_ref=0.00033),
                        ic=dict(k0=8.0, Mt=0.3, c0=float(np.sqrt(11.11111111))), t_over_tau0=7.2, spectra=(),
                        every=2),
}


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def spectrum(U, V, W):
    """Shell-summed kinetic energy spectrum E(k), k = 0 .. N/2, of a periodic
    N^3 velocity field (arrays indexed [z, y, x])."""
    n = U.shape[0]
    kf = np.fft.fftfreq(n, 1.0 / n)
    kr = np.fft.rfftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(kf, kf, kr, indexing="ij")
    shell = np.rint(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)).astype(np.int64)
    # rfft keeps kx >= 0: every other column stands for itself and its conjugate
    weight = np.where((kx == 0) | (kx == n // 2), 1.0, 2.0)
    e = np.zeros(shell.max() + 1)
    for f in (U, V, W):
        fh = np.fft.rfftn(f) / f.size
        e += np.bincount(shell.ravel(), weights=(0.5 * weight * np.abs(fh) ** 2).ravel(), minlength=e.size)
    return e[: n // 2 + 1]


def run(case, steps=None, every=None, spectra=None, seed=0, mu_ref=None):
    c = CASES[case]
    N = c["N"]
    opts = dict(c["opts"], **({} if mu_ref is None else {"mu_ref": mu_ref}))
    g = Solver(hit.options_text(N, **opts), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(hit.hit_initial_condition(N, seed=seed, **c["ic"]))
    every = every or c["every"]
    spectra = set(c["spectra"] if spectra is None else spectra)
    rows, specs = [], {}
    st = g.hit_stats()
    tau0 = st["tau"]
    t_end = c.get("t_over_tau0", 0.0) * tau0
    nsteps = steps or c.get("steps", 0)

    def record(step):
        s = g.hit_stats()
        rows.append([step] + [s[k] for k in COLS])
        if not all(np.isfinite(rows[-1][1:])):
            raise RuntimeError(f"non-finite statistics at step {step}")
        if step in spectra:
            specs[step] = spectrum(g.field("U"), g.field("V"), g.field("W")).tolist()

    record(0)
    done, t_run = 0, 0.0
    while (nsteps and done < nsteps) or (t_end and rows[-1][1] < t_end):
        n = every if not nsteps else min(every, nsteps - done)
        g.sync()
        t0 = time.perf_counter()
        g.step(n)
        g.sync()
        t_run += time.perf_counter() - t0
        done += n
        record(done)
        if done % (20 * every) == 0:
            _say(f"[hit_study] {case} step {done}: t {rows[-1][1]:.4f} KE {0.5 * rows[-1][11]:.6e} "
                 f"eps {rows[-1][5]:.6e} M_t {rows[-1][4]:.4f}")
    g.close()
    res = dict(case=case, N=N, options=opts, seed=seed, initial=c["ic"], steps=done,
               ms_per_step=1e3 * t_run / max(done, 1), columns=["step"] + list(COLS), stats=rows,
               spectra={str(k): v for k, v in sorted(specs.items())})
    if case == "samtaney_d4":
        res["anchor"] = anchor(rows, tau0)
    return res


def anchor(rows, tau0):
    """K/K_0, M_t/M_t0 and rho'/M_t0^2 at the reference's t/tau_0 = 1, 2, 4, 7
    (linear interpolation in t between recorded rows), this run and the
    reference's turbdata.out rows (each normalised by its own first row)."""
    a = np.array(rows)
    t, ke, mt, rp = a[:, 1], 0.5 * a[:, 11], a[:, 4], a[:, 8]
    ref = np.array(SAMTANEY_D4_REF)
    rt0 = ref[0, 1]
    out = []
    for r in ref[1:]:
        x = r[0] / rt0  # the reference row's t / tau_0
        tt = x * tau0
        out.append(dict(t_over_tau0=x,
                        run=dict(K_over_K0=float(np.interp(tt, t, ke) / ke[0]),
                                 Mt_over_Mt0=float(np.interp(tt, t, mt) / mt[0]),
                                 rhoprime_over_Mt0sq=float(np.interp(tt, t, rp) / mt[0] ** 2)),
                        reference=dict(K_over_K0=float(r[10] / ref[0, 10]), Mt_over_Mt0=float(r[3] / ref[0, 3]),
                                       rhoprime_over_Mt0sq=float(r[7] / ref[0, 3] ** 2))))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--case", choices=sorted(CASES), default="les256")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--every", type=int, default=None)
    ap.add_argument("--spectra", default=None, help="comma-separated steps (default: the case's)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    spectra = None if args.spectra is None else [int(s) for s in args.spectra.split(",") if s]
    res = run(args.case, args.steps, args.every, spectra, args.seed)
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
