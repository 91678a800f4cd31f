"""CPU check of the partitioned line-solve plan (host logic of libmc3d.so).

mc3d_plan_tables returns the tables the GPU kernel consumes (chunk LU factors,
spikes, interface-correction rows, stencils).  This test replays the kernel's
arithmetic on them in numpy, line by line, and compares with the oracle, so a
wrong factorisation or coupling is caught before any GPU run.  It also checks
that the C ABI library loads and exports every symbol include/mc3d.h declares.
"""
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from mpi_3dcompact_amd import _lib
    return _lib


def test_library_exports_header_symbols():
    L = _lib().lib()
    hdr = open(os.path.join(ROOT, "include", "mc3d.h")).read()
    names = set(re.findall(r"\b(mc3d_[a-z0-9_]+)\s*\(", hdr))
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(L, n), n


def emulate(t, f):
    """Replay the kernel on one line f using plan tables t."""
    N, W, Pn, H = t["N"], t["W"], t["P"], t["H"]
    nbl, nbr = t["nbl"], t["nbr"]
    ext = np.zeros(N + 2 * H)
    ext[H:H + N] = f
    if t["periodic"]:
        ext[:H] = f[N - H:]
        ext[H + N:] = f[:H]
    rows = t["rows"]
    ys = []
    for c in range(Pn):
        cs, cm = int(t["cstart"][c]), int(t["clen"][c])
        y = np.zeros(16)
        for i in range(cm):
            p = cs + i
            if p < nbl or p >= N - nbr:
                ridx = p if p < nbl else nbl + (p - (N - nbr))
                off, step, wd = int(rows[ridx, 0]), int(rows[ridx, 1]), int(rows[ridx, 2])
                r = 0.0
                for k in range(wd):
                    r = r + rows[ridx, 3 + k] * ext[H + p + off + k * step]
            else:
                r = 0.0
                for k in range(2 * H + 1):
                    r = r + t["interior"][k] * ext[H + p + t["int_off"] + k]
            y[i] = r / t["dd"]
        if W > 0:
            lu = t["lu"][c]
            for i in range(1, 16):
                y[i] -= lu[i, 0] * y[i - 1]
                if W == 2 and i >= 2:
                    y[i] -= lu[i, 1] * y[i - 2]
            for i in range(15, -1, -1):
                x = y[i]
                if i + 1 < 16:
                    x -= lu[i, 2] * y[i + 1]
                if W == 2 and i + 2 < 16:
                    x -= lu[i, 3] * y[i + 2]
                y[i] = x * lu[i, 4]
        ys.append(y)
    if W > 0:
        yhat = np.zeros(t["NI"])
        for c in range(Pn):
            cm = int(t["clen"][c])
            for k in range(W):
                yhat[c * 2 * W + k] = ys[c][k]
                yhat[c * 2 * W + W + k] = ys[c][cm - W + k]
        for c in range(Pn):
            s = t["G"][c] @ yhat
            ys[c] = ys[c] - t["sp"][c] @ s
    out = np.zeros(N)
    for c in range(Pn):
        cs, cm = int(t["cstart"][c]), int(t["clen"][c])
        out[cs:cs + cm] = ys[c][:cm]
    return out


def tables(cls, kind, bc, n, dd=1.0, af=0.0):
    from mpi_3dcompact_amd.ops import plan_tables
    t = plan_tables(cls, kind, bc, n, dd, af)
    t["dd"] = dd if cls == 0 else 1.0
    return t


RNG = np.random.default_rng(7)

DERIVS = [P.PADE6, P.PENTA10, P.CD2]
FILTS = [(P.COMPACT8, 0.49), (P.COMPACT10, 0.49), (P.COMPACT10, -0.49), (P.COMMN5PADE, 0.0),
         (P.COMMN5EXP, 0.0)]


@pytest.mark.parametrize("bc", [P.PERIODIC, P.DIRICHLET])
@pytest.mark.parametrize("n", [11, 17, 32, 48, 64, 96, 256])
@pytest.mark.parametrize("scheme", DERIVS)
def test_deriv_plan_matches_oracle(scheme, n, bc):
    dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
    t = tables(0, scheme, bc, n, dd)
    for _ in range(3):
        f = RNG.normal(size=n)
        ref = P.deriv_line(scheme, bc, f, dd)
        got = emulate(t, f)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("bc", [P.PERIODIC, P.DIRICHLET])
@pytest.mark.parametrize("n", [11, 17, 32, 48, 64, 96, 256])
@pytest.mark.parametrize("kind,af", FILTS)
def test_filter_plan_matches_oracle(kind, af, n, bc):
    t = tables(1, kind, bc, n, 1.0, af)
    for _ in range(3):
        f = RNG.normal(size=n)
        ref = P.filter_line(kind, bc, f, af)
        got = emulate(t, f)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_plan_rejects_short_lines():
    from mpi_3dcompact_amd import Mc3dError
    with pytest.raises(Mc3dError):
        tables(0, P.PADE6, P.PERIODIC, 10)
