"""Pins of the oracle's dynamic Smagorinsky model (test infrastructure
checking test infrastructure) by the invariances the model has by
construction (src/DSMSGS.hpp:124-444; Germano et al. 1991, Moin et al. 1991,
Lilly 1992):

- scaling the velocity by s at constant density and temperature scales
  S_ij, M_ij^(1/2), L_ij^(1/2), N_i, K_i by s: the coefficients C, C_I and
  Pr_t are unchanged, mu_sgs and k_sgs scale by s and tau_kk by s^2 (k_sgs
  and Pr_t only without tau_kk: with it the pressure drops by tau_kk / 2
  ~ s^2 before grad T is taken, T12, which its own check pins);
- a uniform velocity added to a constant-density flow changes no velocity
  gradient and, through hat(const) = const of the linear test filter, no
  Leonard term (L_ij and K_i are differences of filtered products): every
  output is unchanged;
- GLOBAL averaging makes mu_sgs / (rho |S|) = <L M> / <M M> one number and
  XI1 averaging one number per x-line; the coefficient output satisfies
  mu_sgs = rho |S| Delta^2 C and tau_kk = 2 rho |S|^2 Delta^2 C_I.

A plausible restatement slip (a power of |S| or rho dropped, u used where
rho u belongs, a filter applied to the wrong operand, the average taken over
the wrong axis) breaks at least one of these.  The checks are evaluated at
the initial state through the oracle's RHS probe (which runs
calcMuSGSTaukk and calcKSGS as the step does).
"""
import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import pyoracle as P

GAMMA = 1.4
OUT = ("mu_sgs", "taukk", "k_sgs", "C", "CI", "Prt")


def state(N, seed, s=1.0, shift=(0.0, 0.0, 0.0)):
    """Constant density; velocity s * (smooth waves) + shift; pressure varying
    (so the temperature gradient and the k_sgs terms are not zero)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(N) / N, np.arange(N) / N, np.arange(N) / N, indexing="ij")

    def wave(amp):
        f = np.zeros((N, N, N))
        for _ in range(4):
            k = rng.integers(-3, 4, size=3)
            f += np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 2 * np.pi))
        return amp * f / 4

    U, V, W = wave(0.2), wave(0.2), wave(0.2)
    p = 1.0 / GAMMA + wave(0.02)
    return dict(rho=np.ones((N, N, N)), U=s * U + shift[0], V=s * V + shift[1], W=s * W + shift[2], p=p)


def dsm_outputs(N, avg, ic, filt=1, taukk=True):
    cfg = hit.oracle_config(N, fd="PENTA10", filt="COMPACT8", rk="KENRK4", les="DSM", les_avg=avg, mu_ref=1e-4,
                            dsm_filter=filt, use_taukk=taukk, coeff_range=True)
    o = P.Solver(cfg, hit.identity_metrics(N), ic)
    o.rhs()
    return {k: o.field(k) for k in OUT}


def nrel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("avg", ["GLOBAL", "XI1_AVG", "XI3_AVG"])
@pytest.mark.parametrize("filt", [1, 2], ids=["pade", "explicit"])
def test_velocity_scaling(avg, filt):
    """Without tau_kk the temperature is untouched and every output scales;
    with tau_kk (which lowers p by tau_kk / 2 ~ s^2 before grad T, T12) the
    coefficients C, C_I and mu_sgs, tau_kk still do."""
    N = 16
    s = 1.7
    a = dsm_outputs(N, avg, state(N, 1), filt, taukk=False)
    b = dsm_outputs(N, avg, state(N, 1, s=s), filt, taukk=False)
    for k, power in (("C", 0), ("Prt", 0), ("mu_sgs", 1), ("k_sgs", 1)):
        e = nrel(b[k], s ** power * a[k])
        assert e <= 1e-10, (avg, filt, k, e)
    assert np.all(b["taukk"] == 0.0) and np.all(b["CI"] == 0.0)
    a = dsm_outputs(N, avg, state(N, 1), filt)
    b = dsm_outputs(N, avg, state(N, 1, s=s), filt)
    for k, power in (("C", 0), ("CI", 0), ("mu_sgs", 1), ("taukk", 2)):
        e = nrel(b[k], s ** power * a[k])
        assert e <= 1e-10, (avg, filt, k, e)
    # the wrong power fails
    assert nrel(b["mu_sgs"], s ** 2 * a["mu_sgs"]) > 0.1
    assert nrel(b["taukk"], s * a["taukk"]) > 0.1


@pytest.mark.parametrize("avg", ["GLOBAL", "XI1_AVG"])
def test_taukk_enters_the_pressure_before_grad_T(avg):
    """T12 (src/CurvilinearCSolver_Core.cpp:571-594): with DSMUSETAUKK the
    k_sgs of a state equals the k_sgs without tau_kk of the same state with
    p lowered by tau_kk / 2 beforehand (C is unaffected either way)."""
    N = 16
    ic = state(N, 4)
    a = dsm_outputs(N, avg, ic)
    ic2 = dict(ic)
    ic2["p"] = ic["p"] - 0.5 * a["taukk"]
    b = dsm_outputs(N, avg, ic2, taukk=False)
    assert nrel(b["C"], a["C"]) <= 1e-12
    for k in ("k_sgs", "Prt"):
        assert nrel(b[k], a[k]) <= 1e-9, k
    c = dsm_outputs(N, avg, ic, taukk=False)  # grad T of the unmodified pressure differs
    assert nrel(c["k_sgs"], a["k_sgs"]) > 1e-6


@pytest.mark.parametrize("avg", ["GLOBAL", "XI2_AVG"])
def test_uniform_velocity_shift(avg):
    N = 16
    a = dsm_outputs(N, avg, state(N, 2))
    b = dsm_outputs(N, avg, state(N, 2, shift=(0.3, -0.2, 0.15)))
    for k in OUT:
        e = nrel(b[k], a[k])
        assert e <= 1e-9, (avg, k, e)


def test_averaging_structure_and_coefficients():
    """GLOBAL: mu_sgs / (rho |S|) is one number; XI1: one per x-line; and the
    coefficient outputs reproduce mu_sgs and tau_kk."""
    N = 16
    ic = state(N, 3)
    d3 = (1.0 / N) ** 3 * (2 * np.pi) ** 3  # dxi deta dzeta / J of the 2 pi box
    delta2 = np.cbrt(d3) ** 2
    for avg, axes in (("GLOBAL", (0, 1, 2)), ("XI1_AVG", (2,))):
        out = dsm_outputs(N, avg, ic)
        ratio = out["mu_sgs"] / (1.0 * out["C"] * delta2)  # = rho |S| where C != 0
        # C constant over the averaged axes
        C = out["C"]
        spread = np.max(np.abs(C - C.mean(axis=axes, keepdims=True))) / np.max(np.abs(C))
        assert spread <= 1e-13, (avg, spread)
        assert np.all(ratio > 0)
        # tau_kk / (2 rho |S|^2 Delta^2 C_I) = 1 with |S| = mu_sgs / (rho Delta^2 C)
        S = ratio
        tk = 2.0 * S * S * delta2 * out["CI"]
        assert nrel(tk, out["taukk"]) <= 1e-12, avg
