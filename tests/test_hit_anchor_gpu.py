"""Physics anchor against the reference's only published numbers: the
Samtaney D4 decaying-HIT case (tests/hit/Samtaney_D4: 128^3, PADE6, COMPACT10
alpha_f 0.495, LSLDDRK4, CFL 0.5, MU_REF 3.3e-4; rho = 1, p = 11.11/gamma,
M_t 0.3 velocity peaked at k0 = 8) run from this package's seeded field, and
its K/K_0, M_t/M_t0 and rho'/M_t0^2 at t/tau_0 = 1, 2, 4, 7 against
turbdata.out rows 13, 25, 47, 78.  The reference's velocity files are not in
its tree, so the random phases differ and the bar is a decay-shape bar:
measured 1-6% apart (K), 0.5-3% (M_t), 0.4-4% (rho'); held to 10%, 6%, 10%.
"""
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_samtaney_d4_decay_matches_the_published_statistics():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mpi_3dcompact_amd.hit_study import run
    print("case: Samtaney D4 128^3 to t/tau_0 = 7.2, statistics every 10 steps", file=sys.stderr, flush=True)
    res = run("samtaney_d4", every=10)
    s0 = res["stats"][0]
    assert abs(s0[3] / 68.196 - 1) <= 0.05  # initial Re_lambda within 5% of the reference's first row
    for a in res["anchor"]:
        r, ref = a["run"], a["reference"]
        print(f"  t/tau0 {a['t_over_tau0']:.2f}: {r} vs {ref}", file=sys.stderr, flush=True)
        assert abs(r["K_over_K0"] / ref["K_over_K0"] - 1) <= 0.10
        assert abs(r["Mt_over_Mt0"] / ref["Mt_over_Mt0"] - 1) <= 0.06
        assert abs(r["rhoprime_over_Mt0sq"] / ref["rhoprime_over_Mt0sq"] - 1) <= 0.10
