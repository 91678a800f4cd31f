"""Learned-closure kernels on the GPU against the CPU oracle
(oracle/orc_closure.py): inference, gradient, Adam, training, the fused
per-point closure of the solver step and the DSM training-pair sampler.

Tolerances (BF16 operands, FP32 accumulation in a different order than the
oracle's exact sums):
- network outputs: 99% of samples within 1e-5, all within 1e-2 (a BF16
  rounding flip of one activation near a rounding boundary moves a sample by
  up to one BF16 step of that activation's contribution);
- gradient: 1e-3 relative (norm-wise) per parameter block, loss 1e-5;
- Adam with a given gradient: 1 FP32 ulp.
"""
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import orc_closure as C
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _net(seed):
    from mpi_3dcompact_amd.closure import Closure
    p = C.init_params(seed)
    p[-2:] = [0.25, -0.1]
    c = Closure(seed)
    c.set_params(p)
    return c, p


BLOCKS = [(0, 448), (448, 512), (512, 4608), (4608, 4672), (4672, 8768), (8768, 8832), (8832, 8960), (8960, 8962)]


@pytest.mark.parametrize("n", [1, 63, 64, 1000])
def test_inference_matches_oracle(torch, n):
    _say(f"  closure infer n={n}")
    c, p = _net(n)
    x = np.random.default_rng(n).standard_normal((n, 7)).astype(np.float32)
    y = c.infer(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = C.forward(p, x)
    err = np.max(np.abs(y - ref), axis=1)
    _say(f"   max {err.max():.2e}  p99 {np.quantile(err, 0.99):.2e}")
    assert np.quantile(err, 0.99) <= 1e-5 and err.max() <= 1e-2


@pytest.mark.parametrize("n", [5, 64, 300])
def test_gradient_matches_oracle(torch, n):
    _say(f"  closure gradient n={n}")
    c, p = _net(n + 1)
    r = np.random.default_rng(n + 1)
    x = r.standard_normal((n, 7)).astype(np.float32)
    t = r.standard_normal((n, 2)).astype(np.float32)
    g, L = c.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    g = g.cpu().numpy()
    Lr, gr = C.loss_and_grad(p, x, t)
    _say(f"   loss {L:.6e} oracle {Lr:.6e}")
    assert abs(L - Lr) <= 1e-5 * Lr
    for lo, hi in BLOCKS:
        e = np.linalg.norm(g[lo:hi] - gr[lo:hi]) / max(np.linalg.norm(gr[lo:hi]), 1e-30)
        _say(f"   block {lo}:{hi} rel {e:.2e}")
        assert e <= 1e-3, (lo, hi, e)


def test_adam_step_matches_oracle(torch):
    c, p = _net(7)
    r = np.random.default_rng(7)
    ad = C.Adam(C.N_PARAMS)
    for k in range(5):
        g = (r.standard_normal(C.N_PARAMS) * 10.0 ** (k - 3)).astype(np.float32)
        c.adam_step(torch.from_numpy(g).cuda(), 1e-3)
        p = ad.step(p, g)
    got = c.params()
    assert np.max(np.abs(got - p) / np.maximum(np.abs(p), 1e-30)) <= 2.0 ** -23


def test_training_fits_a_smooth_map(torch):
    """Shuffled minibatch Adam on y = (sin x0 + x1 x2, tanh(x3 - x4) + 0.1 x5 x6):
    the loss falls by 10x in 20 epochs and the trained network generalises to
    held-out samples (GPU training, GPU inference)."""
    r = np.random.default_rng(11)
    n = 20000
    x = r.standard_normal((n + 2000, 7)).astype(np.float32)
    y = np.stack([np.sin(x[:, 0]) + x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4]) + 0.1 * x[:, 5] * x[:, 6]],
                 axis=1).astype(np.float32)
    from mpi_3dcompact_amd.closure import Closure
    c = Closure(3)
    X = torch.from_numpy(x[:n]).cuda()
    T = torch.from_numpy(y[:n]).cuda()
    _, L0 = c.gradient(X, T)
    losses = c.train(X, T, epochs=20, batch=512, lr=3e-3, seed=1)
    _say("   losses", " ".join(f"{v:.3e}" for v in [L0] + losses))
    assert losses[-1] < L0 / 10
    assert all(np.isfinite(losses))
    yh = c.infer(torch.from_numpy(x[n:]).cuda()).cpu().numpy()
    held = np.mean((yh - y[n:]) ** 2) / 2
    assert held < 2.0 * losses[-1] + 1e-3


def test_closure_bench_runs(torch, capsys):
    """The training/inference throughput tool runs and reports both rates."""
    from mpi_3dcompact_amd import closure_bench
    assert closure_bench.main(["--samples", "65536", "--epochs", "1", "--batch", "8192"]) == 0
    out = capsys.readouterr().out
    assert "us per batch" in out and "G pairs/s" in out


def test_save_and_load(torch, tmp_path):
    c, p = _net(5)
    c.set_scaling(np.arange(7) * 0.1, np.arange(1, 8) * 0.5, [0.1, 0.2], [0.3, 0.4], -np.arange(1, 8),
                  np.arange(1, 8))
    path = tmp_path / "net.bin"
    c.save(path)
    from mpi_3dcompact_amd.closure import Closure
    d = Closure(99)
    d.load(path)
    assert np.array_equal(d.params(), p)
    s = d.scaling()
    assert np.allclose(s["in_std"], np.arange(1, 8) * 0.5) and np.allclose(s["hi"], np.arange(1, 8))


def _scaled_closure(seed, lo=-1e30, hi=1e30):
    c, p = _net(seed)
    im = np.array([0.0, 0.2, 0.0, -0.01, 0.0, 1.0, -3.0], np.float32)
    isd = np.array([0.1, 0.5, 0.2, 0.1, 0.05, 1.0, 0.5], np.float32)
    om, osd = np.array([0.01, 0.005], np.float32), np.array([0.02, 0.01], np.float32)
    c.set_scaling(im, isd, om, osd, np.full(7, lo), np.full(7, hi))
    return c, p, (im, isd, om, osd)


def _cart(g, shape):
    """Cartesian gradient from the solver's computational derivatives and metrics."""
    J = {k: g.field(k) for k in ("J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32", "J33")}
    A = np.empty(shape + (3, 3))
    for i, u in enumerate("UVW"):
        d = [g.field(f"d{u}{k}") for k in range(3)]
        for j in range(3):
            A[..., i, j] = J[f"J1{j + 1}"] * d[0] + J[f"J2{j + 1}"] * d[1] + J[f"J3{j + 1}"] * d[2]
    return A


def test_solver_closure_matches_oracle(torch):
    """LESMODEL LEARNED on a 32^3 HIT field: mu_sgs and k_sgs of the fused
    kernel against the oracle closure on the solver's own gradient fields."""
    from mpi_3dcompact_amd.solver import Solver
    N = 32
    c, p, (im, isd, om, osd) = _scaled_closure(21)
    g = Solver(hit.options_text(N, fd="PENTA10", filt="COMPACT8", rk="KENRK4", les="LEARNED",
                                closure_fallback=False), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=2))
    g.set_closure(c)
    g.eval_sgs()
    A = _cart(g, (N, N, N))
    vol_J = (1.0 / N) ** 3 / g.field("J")
    f = C.features(A, vol_J)
    x = C.standardise(f.reshape(-1, 7), im, isd)
    y = C.forward(p, x)
    cp = 1.4 / 0.4 * (1.0 / 1.4)
    mu, k = C.closure_outputs(y, om, osd, g.field("rho").ravel(), C.smag(A).ravel(), vol_J.ravel(), cp)
    for name, ref in (("mu_sgs", mu), ("k_sgs", k)):
        got = g.field(name).ravel()
        err = np.abs(got - ref) / np.max(np.abs(ref))
        _say(f"   {name}: max {err.max():.2e} p99 {np.quantile(err, 0.99):.2e}")
        assert np.quantile(err, 0.99) <= 1e-5 and err.max() <= 1e-2


@pytest.mark.parametrize("avg", ["XI1_AVG", "XI3_AVG"])
def test_solver_closure_with_averaged_inputs_matches_oracle(torch, avg):
    """LESAVERAGING applies to the network's inputs as the DSM applies it to its
    contractions: the seven features are averaged along the lines before the
    network, |S| stays local."""
    from mpi_3dcompact_amd.solver import Solver
    N = 32
    c, p, (im, isd, om, osd) = _scaled_closure(22)
    g = Solver(hit.options_text(N, fd="PADE6", filt="COMPACT10", rk="RK4", les="LEARNED", les_avg=avg,
                                closure_fallback=False), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=4))
    g.set_closure(c)
    g.eval_sgs()
    A = _cart(g, (N, N, N))
    vol_J = (1.0 / N) ** 3 / g.field("J")
    f = C.averaged_features(A, vol_J, avg)
    x = C.standardise(f.reshape(-1, 7), im, isd)
    y = C.forward(p, x)
    cp = 1.4 / 0.4 * (1.0 / 1.4)
    mu, k = C.closure_outputs(y, om, osd, g.field("rho").ravel(), C.smag(A).ravel(), vol_J.ravel(), cp)
    for name, ref in (("mu_sgs", mu), ("k_sgs", k)):
        got = g.field(name).ravel()
        err = np.abs(got - ref) / np.max(np.abs(ref))
        _say(f"   {avg} {name}: max {err.max():.2e} p99 {np.quantile(err, 0.99):.2e}")
        assert np.quantile(err, 0.99) <= 1e-5 and err.max() <= 1e-2
    # a coefficient constant along the lines: mu_sgs / (rho Delta^2 |S|) is
    axis = 2 if avg == "XI1_AVG" else 0
    coef = (g.field("mu_sgs") / (g.field("rho") * np.cbrt(vol_J) ** 2 * C.smag(A)))
    assert np.allclose(coef, np.take(coef, [0], axis=axis), rtol=1e-12, atol=0)


def test_point_features_ignore_the_averaging(torch):
    """CLOSUREFEATURES POINT: under LESAVERAGING LOCAL the network reads the
    features at each point, so mu_sgs and k_sgs equal those of LESAVERAGING
    NONE bit for bit (the same kernel on the same values); LESAVERAGING still
    sets the fallback DSM's averaging."""
    from mpi_3dcompact_amd.solver import Solver
    N = 32
    c, _, _ = _scaled_closure(24)
    out = {}
    for avg, feats in (("NONE", None), ("LOCAL", "POINT")):
        g = Solver(hit.options_text(N, fd="PADE6", filt="COMPACT10", rk="RK4", les="LEARNED", les_avg=avg,
                                    closure_fallback=False, closure_features=feats), (N, N, N))
        g.set_metrics(hit.identity_metrics(N))
        g.set_initial(hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=5))
        g.set_closure(c)
        g.eval_sgs()
        out[avg] = (g.field("mu_sgs"), g.field("k_sgs"))
        g.close()
    assert np.array_equal(out["NONE"][0], out["LOCAL"][0]) and np.array_equal(out["NONE"][1], out["LOCAL"][1])


def test_solver_closure_steps_and_falls_back(torch):
    """A narrow trusted range sends the stages to the DSM fallback; with the
    range wide open no stage falls back; with a fallback fraction of 1 the
    out-of-range points are only clamped.  The runs stay finite."""
    from mpi_3dcompact_amd.solver import Solver
    N = 32
    for lo, hi, frac, expect in ((-1e30, 1e30, None, False), (-0.5, 0.5, None, True), (-0.5, 0.5, 1.0, False)):
        c, _, _ = _scaled_closure(23, lo, hi)
        g = Solver(hit.options_text(N, fd="PADE6", filt="COMPACT10", rk="TVDRK3", les="LEARNED",
                                    les_avg="LOCAL", closure_fallback_fraction=frac), (N, N, N))
        g.set_metrics(hit.identity_metrics(N))
        g.set_initial(hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=3))
        g.set_closure(c)
        g.step(3)
        st = g.closure_stats()
        _say(f"   range [{lo}, {hi}] fraction {frac}: {st}")
        assert st["evaluations"] == 9
        assert (st["fallback_stages"] > 0) == expect
        assert (st["outside_points"] > 0) == (lo > -1e30)
        for name in ("rho", "rhoE", "mu_sgs"):
            assert np.all(np.isfinite(g.field(name)))


@pytest.mark.parametrize("avg", ["GLOBAL", "XI1_AVG"])
def test_dsm_sampling_matches_its_fields(torch, avg):
    """Training pairs of a DSM state: features equal the oracle's on the DSM's
    own gradients, averaged as the DSM averages (LESAVERAGING), targets equal
    its C and C / Pr_t, and the state is unchanged."""
    from mpi_3dcompact_amd.solver import Solver
    N = 32
    g = Solver(hit.options_text(N, fd="PENTA10", filt="COMPACT8", rk="KENRK4", les="DSM", les_avg=avg,
                                coeff_range=True), (N, N, N))
    g.set_metrics(hit.identity_metrics(N))
    g.set_initial(hit.hit_initial_condition(N, k0=4.0, Mt=0.3, seed=5))
    g.step(2)
    before = {k: g.field(k) for k in ("rho", "rhoE", "p", "T", "mu")}
    count, off, stride = 2000, 17, 13
    X = torch.empty((count, 7), dtype=torch.float32, device="cuda")
    T = torch.empty((count, 2), dtype=torch.float32, device="cuda")
    g.sample_closure(X, T, count, off, stride)
    for k, v in before.items():
        assert np.array_equal(g.field(k), v), k
    A = _cart(g, (N, N, N))
    vol_J = (1.0 / N) ** 3 / g.field("J")
    pts = off + stride * np.arange(count)
    # the global mean is summed in another order on the GPU (block partial sums)
    f = C.averaged_features(A, vol_J, avg).reshape(-1, 7)[pts].astype(np.float32)
    Xg = X.cpu().numpy()
    assert np.max(np.abs(Xg - f) / np.maximum(np.abs(f), 1e-3)) <= 1e-6
    Cf, Pf = g.field("C").ravel()[pts], g.field("Prt").ravel()[pts]
    Tg = T.cpu().numpy()
    assert np.array_equal(Tg[:, 0], Cf.astype(np.float32))
    assert np.array_equal(Tg[:, 1], (Cf / Pf).astype(np.float32))


def test_study_pipeline_small(torch, tmp_path):
    """The whole learned-closure pipeline at 128^3 (the LOCAL-averaged DSM is
    unstable on coarser grids): DSM sampling, standardisation, GPU training,
    online runs.  What the network gains over the method without it (constant
    training-mean coefficients): its held-out error on the DSM's coefficients
    is below the constant predictor's (standardised MSE 1)."""
    import json
    from mpi_3dcompact_amd import closure_study
    out = tmp_path / "study.json"
    closure_study.main(["--N", "128", "--gen-steps", "40", "--snapshots", "4", "--samples", "40000",
                        "--epochs", "8", "--batch", "1024", "--online-steps", "40", "--stats-every", "10",
                        "--out", str(out)])
    res = json.loads(out.read_text())
    _say("   held-out MSE", res["heldout_mse"], "errors", res["learned_vs_dsm"], res["constant_vs_dsm"],
         "ms/step", res["ms_per_step"], "fallback", res["fallback"])
    assert res["samples"] > 20000
    assert all(np.isfinite(res["epoch_losses"])) and res["epoch_losses"][-1] < res["epoch_losses"][0]
    assert min(res["heldout_mse"]) < 0.9
    for k in ("learned_vs_dsm", "constant_vs_dsm", "nomodel_vs_dsm"):
        assert np.isfinite(res[k]["ke_max_rel"]) and np.isfinite(res[k]["dissipation_max_rel"])
    assert res["fallback"]["evaluations"] == 40 * 5
