"""CPU check of the partitioned line-solve plan (host logic of libmc3d.so).

mc3d_plan_tables returns the tables the GPU kernels consume (chunk LU factors,
spikes, interface-correction rows, stencils, dense end rows).  This test replays
the arithmetic of both kernel families (the LDS-tiled kernel for any chunking
and the fast kernels for evenly chunked lines) on them in numpy, line by line,
and compares with the oracle, so a wrong factorisation, coupling or table
layout is caught before any GPU run.  It also checks
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
        MM = t["MM"]
        y = np.zeros(MM)
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
            for i in range(1, MM):
                y[i] -= lu[i, 0] * y[i - 1]
                if W == 2 and i >= 2:
                    y[i] -= lu[i, 1] * y[i - 2]
            for i in range(MM - 1, -1, -1):
                x = y[i]
                if i + 1 < MM:
                    x -= lu[i, 2] * y[i + 1]
                if W == 2 and i + 2 < MM:
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


def _lu_solve(lu, y, W):
    MM = len(y)
    for i in range(1, MM):
        y[i] -= lu[i, 0] * y[i - 1]
        if W == 2 and i >= 2:
            y[i] -= lu[i, 1] * y[i - 2]
    for i in range(MM - 1, -1, -1):
        x = y[i]
        if i + 1 < MM:
            x -= lu[i, 2] * y[i + 1]
        if W == 2 and i + 2 < MM:
            x -= lu[i, 3] * y[i + 2]
        y[i] = x * lu[i, 4]


def _dense_row(d, seg):
    """Dense special row d = {order, w[K]} over the K = len(seg) points seg."""
    r = 0.0
    K = len(seg)
    order = range(K) if d[0] > 0 else range(K - 1, -1, -1)
    for k in order:
        r = r + d[1 + k] * seg[k]
    return r


def ragged(t):
    """MM-row chunks (16 or 32) and one short second chunk: the fast kernels'
    ragged layout."""
    clen, MM, P = list(t["clen"]), t["MM"], t["P"]
    return (t["N"] != P * MM and MM in (16, 32) and P >= 3 and clen[1] < MM
            and all(c == MM for i, c in enumerate(clen) if i != 1))


def fast_eligible(t):
    if t["periodic"]:
        rows_ok = t["nbl"] + t["nbr"] == 0 or (t["dense_ok"] and t["W"] == 0 and t["H"] == 1)
    else:
        rows_ok = t["dense_ok"] and t["nbl"] == t["H"] == t["nbr"]
    return (t["N"] == t["P"] * t["MM"] or ragged(t)) and rows_ok


def emulate_fast(t, f):
    """Replay the fast kernels (evenly chunked lines, or ragged ones: MM-row
    chunks and a short second chunk) on one line f: an MM + 2H register window
    per chunk, dense end rows, multiplication by an exact 1/dd, chunk 0 factors
    for every chunk of an evenly chunked periodic line and chunk 0's interface
    rows rotated (chunk-own rows otherwise), the short chunk's last interface
    values from its own last rows, and only the interface blocks within gb
    chunks of each chunk."""
    N, W, Pn, H, MM, NI = t["N"], t["W"], t["P"], t["H"], t["MM"], t["NI"]
    nbl, nbr, per = t["nbl"], t["nbr"], t["periodic"]
    rg = ragged(t)
    own = not per or rg  # chunk-own factor and interface rows
    dd = t["dd"]
    m, ex = np.frexp(dd)
    inv = 1.0 / dd if (t["divide"] and m == 0.5) else 0.0
    ys = []
    for c in range(Pn):
        cs, ln = int(t["cstart"][c]), int(t["clen"][c])
        idx = np.arange(cs - H, cs + MM + H)
        if per:
            w = f[idx % N]
        else:
            w = np.where((idx >= 0) & (idx < N), f[np.clip(idx, 0, N - 1)], 0.0)
        y = np.zeros(MM)
        K = 2 * H + 1
        for i in range(MM):
            if c == 0 and i < nbl:
                r = _dense_row(t["drows"][i], w[i:i + K] if per else w[H:H + K])
            elif c == Pn - 1 and i >= MM - nbr:
                r = _dense_row(t["drows"][nbl + i - (MM - nbr)], w[i:i + K] if per else w[H + MM - K:H + MM])
            else:
                r = 0.0
                for k in range(2 * H + 1):
                    r = r + t["interior"][k] * w[i + k]
            if inv != 0.0:
                r = r * inv
            elif t["divide"]:
                r = r / dd
            y[i] = r
        if W > 0:
            _lu_solve(t["lu"][c if own else 0], y, W)
        ys.append(y)
    lens = [int(v) for v in t["clen"]]
    if W > 0:
        yhat = np.concatenate([np.concatenate([y[:W], y[ln - W:ln]]) for y, ln in zip(ys, lens)])
        for c in range(Pn):
            if own:
                G = t["G"][c].copy()
            else:
                G = np.roll(t["G"][0], c * 2 * W, axis=1)
            dist = np.abs(np.arange(Pn) - c)
            if per:
                dist = np.minimum(dist, Pn - dist)
            G = G * np.repeat(dist <= t["gb"], 2 * W)[None, :]
            s = G @ yhat
            ys[c] = ys[c] - t["sp"][c if own else 0] @ s
    return np.concatenate([y[:ln] for y, ln in zip(ys, lens)])


def tables(cls, kind, bc, n, dd=1.0, af=0.0, mm=0):
    from mpi_3dcompact_amd.ops import plan_tables
    t = plan_tables(cls, kind, bc, n, dd, af, mm)
    t["dd"] = dd if cls == 0 else 1.0
    t["divide"] = cls == 0
    return t


RNG = np.random.default_rng(7)

DERIVS = [P.PADE6, P.PENTA10, P.CD2]
FILTS = [(P.COMPACT8, 0.49), (P.COMPACT10, 0.49), (P.COMPACT10, -0.49), (P.COMMN5PADE, 0.0),
         (P.COMMN5EXP, 0.0)]


def _check(t, ref_fn, n):
    for _ in range(3):
        f = RNG.normal(size=n)
        ref = ref_fn(f)
        tol = 1e-12 * max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(emulate(t, f) - ref)) <= tol
        if fast_eligible(t):
            assert np.max(np.abs(emulate_fast(t, f) - ref)) <= tol


@pytest.mark.parametrize("mm", [0, 16])
@pytest.mark.parametrize("bc", [P.PERIODIC, P.DIRICHLET])
@pytest.mark.parametrize("n", [11, 17, 32, 36, 40, 48, 56, 64, 72, 96, 131, 150, 200, 256, 310])
@pytest.mark.parametrize("scheme", DERIVS)
def test_deriv_plan_matches_oracle(scheme, n, bc, mm):
    dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
    t = tables(0, scheme, bc, n, dd, mm=mm)
    _check(t, lambda f: P.deriv_line(scheme, bc, f, dd), n)


@pytest.mark.parametrize("mm", [0, 16])
@pytest.mark.parametrize("bc", [P.PERIODIC, P.DIRICHLET])
@pytest.mark.parametrize("n", [11, 17, 32, 36, 40, 48, 56, 64, 72, 96, 131, 150, 200, 256, 310])
@pytest.mark.parametrize("kind,af", FILTS)
def test_filter_plan_matches_oracle(kind, af, n, bc, mm):
    t = tables(1, kind, bc, n, 1.0, af, mm=mm)
    _check(t, lambda f: P.filter_line(kind, bc, f, af), n)


def test_interface_band():
    """The interface rows G_c decay geometrically away from chunk c; the fast
    kernels multiply only the blocks within gb chunks, those beyond being below
    1e-17 of G_c's largest entry.  Band widths of 256-point lines, and the
    banded replay equal to the full one to 1e-15 of the result."""
    cases = {(P.PENTA10, 16): 5, (P.PENTA10, 32): 3, (P.PADE6, 16): 3, (P.PADE6, 32): 2}
    for (scheme, mm), gb in cases.items():
        for bc in (P.PERIODIC, P.DIRICHLET):
            dd = 1 / 256 if bc == P.PERIODIC else 1 / 255
            t = tables(0, scheme, bc, 256, dd, mm=mm)
            assert t["gb"] == gb, (scheme, mm, bc, t["gb"])
            f = RNG.normal(size=256)
            full = emulate_fast(dict(t, gb=t["P"]), f)
            band = emulate_fast(t, f)
            assert np.max(np.abs(band - full)) <= 1e-15 * np.max(np.abs(full)), (scheme, mm, bc)
    # Compact8 at alpha_f 0.49 decays slowly (about 30x per 16 rows): no band
    t = tables(1, P.COMPACT8, P.PERIODIC, 256, 1.0, 0.49, mm=16)
    assert 2 * t["gb"] + 1 >= t["P"]


def test_chunking_choice():
    """32-row chunks when a line splits evenly into at least two of them."""
    assert tables(0, P.PENTA10, P.PERIODIC, 256, 1 / 256)["MM"] == 32
    assert tables(0, P.PENTA10, P.PERIODIC, 256, 1 / 256)["P"] == 8
    assert tables(0, P.PENTA10, P.PERIODIC, 48, 1 / 48)["MM"] == 16
    assert tables(0, P.PENTA10, P.PERIODIC, 32, 1 / 32)["MM"] == 16
    assert tables(0, P.PADE6, P.DIRICHLET, 17, 1 / 16)["MM"] == 16
    # lengths that are not multiples of 16: the largest of 28, 24, 20 rows (and
    # 12, 8 on periodic lines) that splits the line into 2..16 equal chunks
    for n, bc, mm, chunks in ((200, P.PERIODIC, 20, 10), (200, P.DIRICHLET, 20, 10), (56, P.DIRICHLET, 28, 2),
                              (72, P.PERIODIC, 24, 3), (36, P.PERIODIC, 12, 3), (24, P.PERIODIC, 12, 2),
                              (88, P.PERIODIC, 8, 11), (104, P.PERIODIC, 8, 13)):
        t = tables(0, P.PENTA10, bc, n, 1.0 / n)
        assert (t["MM"], t["P"]) == (mm, chunks), (n, bc)
    # no such divisor, or a Dirichlet line that only 12 rows would split: 16-row
    # chunks (32 above 256 points) and one short second chunk (the ragged
    # layout of the fast kernels) when there are at least three chunks and the
    # short one holds 2W rows; else chunks of equal length and a shorter last one
    for n, bc, mm, lens in ((150, P.PERIODIC, 16, [16, 6] + [16] * 8), (36, P.DIRICHLET, 16, [16, 4, 16]),
                            (131, P.DIRICHLET, 16, [16, 3] + [16] * 7), (310, P.PERIODIC, 32, [32, 22] + [32] * 8),
                            (272, P.DIRICHLET, 32, [32, 16] + [32] * 7)):
        t = tables(0, P.PENTA10 if n != 131 else P.PADE6, bc, n, 1.0 / n)
        assert t["MM"] == mm and list(t["clen"]) == lens, (n, list(t["clen"]))
        assert list(t["cstart"]) == list(np.cumsum([0] + lens[:-1])) and ragged(t) and fast_eligible(t)
    t = tables(0, P.PENTA10, P.DIRICHLET, 131, 1.0 / 131)  # a 3-row short chunk cannot hold 2W = 4 rows
    assert t["MM"] == 16 and list(t["clen"]) == [15] * 8 + [11] and not fast_eligible(t)
    assert not ragged(tables(0, P.PENTA10, P.PERIODIC, 20, 1 / 20))  # two chunks only
    from mpi_3dcompact_amd._lib import Mc3dError
    with pytest.raises(Mc3dError):
        tables(0, P.PENTA10, P.PERIODIC, 100, 1 / 100, mm=10)


DENSE_CASES = [(0, P.PADE6, P.DIRICHLET), (0, P.PENTA10, P.DIRICHLET), (0, P.CD2, P.DIRICHLET),
               (0, P.CD2, P.PERIODIC), (1, P.COMPACT8, P.DIRICHLET), (1, P.COMPACT10, P.DIRICHLET),
               (1, P.COMMN5PADE, P.DIRICHLET), (1, P.COMMN5EXP, P.DIRICHLET)]


@pytest.mark.parametrize("cls,kind,bc", DENSE_CASES)
@pytest.mark.parametrize("n", [32, 64, 256])
def test_dense_rows_match_row_records_bitwise(cls, kind, bc, n):
    """The dense special rows (K = 2H + 1 taps over the window points the
    row's position gives it) sum the same taps in the same order as the row
    records (zero taps add exact zeros), so both kernel families produce the
    same bits for the explicit stage at the line ends and, for periodic CD2,
    at its sign-flipped wrap rows."""
    t = tables(cls, kind, bc, n, 1.0 / (n - 1), 0.49 if cls == 1 else 0.0)
    assert t["dense_ok"]
    f = RNG.normal(size=n)
    rows, nbl, nbr, H = t["rows"], t["nbl"], t["nbr"], t["H"]
    K = 2 * H + 1
    assert nbl + nbr > 0
    for ridx in range(nbl + nbr):
        p = ridx if ridx < nbl else n - nbr + (ridx - nbl)
        off, step, wd = int(rows[ridx, 0]), int(rows[ridx, 1]), int(rows[ridx, 2])
        r = 0.0
        for k in range(wd):
            r = r + rows[ridx, 3 + k] * f[(p + off + k * step) % n]
        if bc == P.PERIODIC:
            seg = f[(p - H + np.arange(K)) % n]
        else:
            seg = f[:K] if ridx < nbl else f[n - K:]
        assert _dense_row(t["drows"][ridx], seg) == r
        assert not np.any(t["drows"][ridx][1 + K:])


@pytest.mark.parametrize("kind,cls,af", [(P.PENTA10, 0, 0.0), (P.PADE6, 0, 0.0), (P.COMPACT8, 1, 0.49),
                                         (P.COMMN5PADE, 1, 0.0)])
def test_periodic_uniform_chunks_share_factors(kind, cls, af):
    """Evenly chunked periodic lines: every chunk has chunk 0's LU factors and
    spikes, and its interface rows are chunk 0's rotated (the fast kernels
    rely on both)."""
    t = tables(cls, kind, P.PERIODIC, 256, 1 / 256, af)
    W2 = 2 * t["W"]
    for c in range(1, t["P"]):
        assert np.array_equal(t["lu"][c], t["lu"][0])
        assert np.array_equal(t["sp"][c], t["sp"][0])
        assert np.array_equal(t["G"][c], np.roll(t["G"][0], c * W2, axis=1))


@pytest.mark.parametrize("kind,cls,af", [(P.PENTA10, 0, 0.0), (P.PADE6, 0, 0.0), (P.COMPACT8, 1, 0.49),
                                         (P.COMPACT10, 1, -0.49), (P.COMMN5PADE, 1, 0.0)])
@pytest.mark.parametrize("n", [544, 576, 960])
def test_dirichlet_interior_chunks_share_factors(kind, cls, af, n):
    """Evenly chunked Dirichlet lines of 32-row chunks: chunks 1 .. P - 2 hold
    bitwise the same LU factors and spikes (the two-chunk y/z kernel gives
    every interior chunk chunk 1's), the end chunks their own (the matrix's
    boundary rows), and the interface rows differ from chunk to chunk."""
    t = tables(cls, kind, P.DIRICHLET, n, 1 / (n - 1), af, 32)
    assert t["P"] == n // 32
    for c in range(2, t["P"] - 1):
        assert np.array_equal(t["lu"][c], t["lu"][1])
        assert np.array_equal(t["sp"][c], t["sp"][1])
    for c in (0, t["P"] - 1):
        assert not np.array_equal(t["lu"][c], t["lu"][1])
    assert not np.array_equal(t["G"][2], np.roll(t["G"][1], 2 * t["W"], axis=1))


@pytest.mark.parametrize("kind,cls,af", [(P.PENTA10, 0, 0.0), (P.PADE6, 0, 0.0), (P.COMPACT8, 1, 0.49),
                                         (P.COMMN5PADE, 1, 0.0)])
@pytest.mark.parametrize("n", [600, 650, 906])
def test_dirichlet_ragged_interior_chunks_share_factors(kind, cls, af, n):
    """Ragged Dirichlet lines of 32-row chunks (the short chunk second):
    chunks 2 .. P - 2 hold bitwise chunk 2's LU factors and spikes (the
    two-chunk y/z kernel gives them all chunk 2's); the short chunk and the
    end chunks their own."""
    t = tables(cls, kind, P.DIRICHLET, n, 1 / (n - 1), af, 32)
    Pn = t["P"]
    assert Pn == n // 32 + 1 and list(t["clen"][:3]) == [32, n % 32, 32]
    for c in range(3, Pn - 1):
        assert np.array_equal(t["lu"][c], t["lu"][2])
        assert np.array_equal(t["sp"][c], t["sp"][2])
    for c in (0, 1, Pn - 1):
        assert not np.array_equal(t["lu"][c], t["lu"][2])


def test_plan_rejects_short_lines():
    from mpi_3dcompact_amd import Mc3dError
    with pytest.raises(Mc3dError):
        tables(0, P.PADE6, P.PERIODIC, 10)
