"""The energy spectrum of hit_study (host numpy; no GPU): Parseval's sum and
single Fourier modes landing in their shell."""
import numpy as np

from mpi_3dcompact_amd import hit
from mpi_3dcompact_amd.hit_study import anchor, spectrum


def test_spectrum_sums_to_the_kinetic_energy():
    ic = hit.hit_initial_condition(32, k0=4.0, Mt=0.3, seed=3)
    e = spectrum(ic["U"], ic["V"], ic["W"])
    ke = 0.5 * np.mean(ic["U"] ** 2 + ic["V"] ** 2 + ic["W"] ** 2)
    assert abs(e.sum() - ke) <= 1e-12 * ke
    assert np.argmax(e) == 4  # E(k) ~ k^4 exp(-2 (k/4)^2) peaks at k0


def test_single_modes_fall_in_their_shell():
    n = 24
    t = 2 * np.pi * np.arange(n) / n
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    zero = np.zeros((n, n, n))
    # u = cos(3x): energy 1/4 at |k| = 3; v = sin(2y + 2z): 1/4 at |k| = sqrt(8) -> shell 3
    e = spectrum(np.cos(3 * x), np.sin(2 * y + 2 * z), zero)
    assert np.allclose(e[3], 0.5, atol=1e-14) and np.allclose(np.delete(e, 3), 0.0, atol=1e-14)
    # the Nyquist column counts once
    e = spectrum(np.cos(12 * x), zero, zero)
    assert np.allclose(e[12], 0.5, atol=1e-14)


def test_anchor_normalises_by_the_first_row():
    # a run whose columns equal the reference rows' gives the reference's ratios
    from mpi_3dcompact_amd.hit_study import SAMTANEY_D4_REF
    rows = [[i] + list(r) for i, r in enumerate(SAMTANEY_D4_REF)]
    out = anchor(rows, SAMTANEY_D4_REF[0][1])
    for a in out:
        for k in ("K_over_K0", "Mt_over_Mt0", "rhoprime_over_Mt0sq"):
            assert abs(a["run"][k] - a["reference"][k]) <= 1e-12 * abs(a["reference"][k])
