"""GPU parity of the derivative / filter line operators against the oracle.

Tolerance: normwise relative error max|gpu - oracle| / max|oracle| <= 1e-12 per
application (single derivative or filter application, FP64).
"""
import itertools
import sys

import numpy as np
import pytest

from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def field(shape, seed):
    rng = np.random.default_rng(seed)
    Nz, Ny, Nx = shape
    z, y, x = np.meshgrid(np.arange(Nz) / Nz, np.arange(Ny) / Ny, np.arange(Nx) / Nx, indexing="ij")
    f = 1.0 + 0.01 * rng.normal(size=shape)
    for _ in range(6):
        k = rng.integers(-4, 5, size=3)
        f += 0.3 * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 6.28))
    return f


SHAPES = [(12, 20, 16), (32, 48, 64), (64, 64, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("scheme,bc", list(itertools.product([P.PADE6, P.PENTA10, P.CD2],
                                                               [P.PERIODIC, P.DIRICHLET])))
def test_deriv_all_directions(torch_cuda, scheme, bc, shape):
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 1)
    fin = torch.from_numpy(f).cuda()
    for d in range(3):
        n = dims[d]
        dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
        _say(f"deriv scheme={scheme} bc={bc} shape={shape} dir={d}")
        op = LineOp.deriv(scheme, d, bc, dims, dd)
        out = torch.empty_like(fin)
        op.apply([fin], [out])
        torch.cuda.synchronize()
        ref = P.deriv_field(scheme, d, bc, f, dd)
        assert relerr(out.cpu().numpy(), ref) <= TOL, (scheme, bc, d)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind,af", [(P.COMPACT8, 0.49), (P.COMPACT10, 0.49), (P.COMPACT10, -0.49),
                                     (P.COMMN5PADE, 0.0), (P.COMMN5EXP, 0.0)])
@pytest.mark.parametrize("bc3", [(0, 0, 0), (0, 1, 1), (1, 0, 1)])
def test_filter_all_directions(torch_cuda, kind, af, bc3, shape):
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 2)
    fin = torch.from_numpy(f).cuda()
    for d in range(3):
        _say(f"filter kind={kind} af={af} bc3={bc3} shape={shape} dir={d}")
        op = LineOp.filter(kind, af, d, bc3, dims)
        out = torch.empty_like(fin)
        op.apply([fin], [out])
        torch.cuda.synchronize()
        ref = P.filter_field(kind, af, d, list(bc3), f)
        assert relerr(out.cpu().numpy(), ref) <= TOL, (kind, bc3, d)


# ragged lengths: x 150 = 16 + 6 + 8 x 16 (periodic) / 150; y 36 = 16 + 4 + 16;
# z 131 = 16 + 3 + 7 x 16 (three-row short chunk: W <= 1 only); second shape
# x 310 = 32 + 22 + 8 x 32
RAGGED_SHAPES = [(131, 36, 150), (12, 20, 310)]


@pytest.mark.parametrize("shape", RAGGED_SHAPES)
@pytest.mark.parametrize("scheme,bc", list(itertools.product([P.PADE6, P.PENTA10, P.CD2],
                                                               [P.PERIODIC, P.DIRICHLET])))
def test_deriv_ragged_lines(torch_cuda, scheme, bc, shape):
    """Lines that no chunk length splits evenly: 16- or 32-row chunks and one
    short second chunk on the fast kernels, against the oracle at 1e-12."""
    from mpi_3dcompact_amd.ops import LineOp, plan_tables
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 3)
    fin = torch.from_numpy(f).cuda()
    for d in range(3):
        n = dims[d]
        dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
        op = LineOp.deriv(scheme, d, bc, dims, dd)
        kname = op.kernel()
        _say(f"ragged deriv scheme={scheme} bc={bc} shape={shape} dir={d} kernel={kname}")
        t = plan_tables(0, scheme, bc, n, dd)
        if n % t["MM"] and t["P"] >= 3 and n - (t["P"] - 1) * t["MM"] >= max(2 * t["W"], 1) and n <= 512:
            assert kname.startswith("lineop_contig" if d == 0 else "lineop_strided"), kname
        out = torch.empty_like(fin)
        op.apply([fin], [out])
        torch.cuda.synchronize()
        ref = P.deriv_field(scheme, d, bc, f, dd)
        assert relerr(out.cpu().numpy(), ref) <= TOL, (scheme, bc, d)


@pytest.mark.parametrize("shape", RAGGED_SHAPES)
@pytest.mark.parametrize("kind,af", [(P.COMPACT8, 0.49), (P.COMPACT10, -0.49), (P.COMMN5PADE, 0.0),
                                     (P.COMMN5EXP, 0.0)])
@pytest.mark.parametrize("bc3", [(0, 0, 0), (1, 1, 1)])
def test_filter_ragged_lines(torch_cuda, kind, af, bc3, shape):
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 4)
    fin = torch.from_numpy(f).cuda()
    for d in range(3):
        op = LineOp.filter(kind, af, d, bc3, dims)
        _say(f"ragged filter kind={kind} bc3={bc3} shape={shape} dir={d} kernel={op.kernel()}")
        out = torch.empty_like(fin)
        op.apply([fin], [out])
        torch.cuda.synchronize()
        ref = P.filter_field(kind, af, d, list(bc3), f)
        assert relerr(out.cpu().numpy(), ref) <= TOL, (kind, bc3, d)
        fl = torch.from_numpy(f).cuda()  # in place
        op.apply([fl], [fl])
        torch.cuda.synchronize()
        assert relerr(fl.cpu().numpy(), ref) <= TOL, (kind, bc3, d, "in place")


@pytest.mark.parametrize("nx", [310, 384, 512, 544, 640, 960])
@pytest.mark.parametrize("bc", [P.PERIODIC, P.DIRICHLET])
def test_long_x_lines_in_phases(torch_cuda, nx, bc):
    """x-lines whose 64-line tile exceeds LDS (banded Dirichlet operators, and
    ragged periodic ones): the chunks are staged in groups through one tile
    (lineop_contig_ph); banded Dirichlet lines of 17 to 30 chunks (544 = 17 x
    32 with an idle upper half-wave, 640, 960 = 30 x 32) in 32-line tiles two
    chunks to a wave, staged in groups of waves (lineop_contig_pair), 216
    lines (a partial last workgroup); derivatives and filters against the
    oracle, pass-through lines and in-place filtering included."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    shape = (12, 18, nx)
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 9)
    fin = torch.from_numpy(f).cuda()
    out = torch.empty_like(fin)
    dd = 1.0 / nx if bc == P.PERIODIC else 1.0 / (nx - 1)
    phased = bc == P.DIRICHLET or nx == 310
    expect = "lineop_contig_pair<" if bc == P.DIRICHLET and nx > 512 else "lineop_contig_ph<" if phased else None
    for scheme in (P.PADE6, P.PENTA10):
        op = LineOp.deriv(scheme, 0, bc, dims, dd)
        kname = op.kernel()
        _say(f"long x deriv nx={nx} bc={bc} scheme={scheme} kernel={kname}")
        if expect:
            assert kname.startswith(expect), kname
        op.apply([fin], [out])
        torch.cuda.synchronize()
        assert relerr(out.cpu().numpy(), P.deriv_field(scheme, 0, bc, f, dd)) <= TOL, (scheme, nx, bc)
    bc3 = (bc, 1, 1)
    for kind, af in ((P.COMPACT8, 0.49), (P.COMPACT10, -0.49), (P.COMMN5PADE, 0.0)):
        op = LineOp.filter(kind, af, 0, bc3, dims)
        kname = op.kernel()
        _say(f"long x filter nx={nx} bc3={bc3} kind={kind} kernel={kname}")
        if expect:
            assert kname.startswith(expect), kname
        ref = P.filter_field(kind, af, 0, list(bc3), f)
        op.apply([fin], [out])
        torch.cuda.synchronize()
        assert relerr(out.cpu().numpy(), ref) <= TOL, (kind, nx, bc)
        fl = torch.from_numpy(f).cuda()
        op.apply([fl], [fl])
        torch.cuda.synchronize()
        assert relerr(fl.cpu().numpy(), ref) <= TOL, (kind, nx, bc, "in place")


@pytest.mark.parametrize("shape", [(576, 40, 12), (544, 13, 11), (960, 12, 16), (650, 12, 16), (560, 12, 11),
                                   (906, 10, 10), (1024, 12, 16)])
def test_long_yz_lines_in_rounds(torch_cuda, shape):
    """z-lines of more than 16 chunks: evenly chunked lines two chunks to a
    wave (lineop_strided_pair): periodic ones without special rows up to 32
    chunks (576 = 18 x 32, 544 = 17 x 32 with the last wave's upper half
    idle, 1024 = 32 x 32), Dirichlet ones and CD2 up to 30 (a wave for each
    end chunk; 960 = 30 x 32 fills 16 waves), ragged Dirichlet ones up to 29
    (650 = 32 + 10 + 19 x 32, the short chunk a wave of its own; 560 = 32 +
    16 + 16 x 32 with the last wave's upper half idle; 906 = 29 chunks in 16
    waves); the rest
    (periodic 650 = 16 + 10 + 39 x 16, Dirichlet and CD2 at 1024) in 16-row
    chunks, 16 chunk waves taking them in rounds (lineop_strided_multi); every
    operator against the oracle, in place too."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 12)
    fin = torch.from_numpy(f).cuda()
    out = torch.empty_like(fin)
    for bc in (P.PERIODIC, P.DIRICHLET):
        dd = 1.0 / Nz if bc == P.PERIODIC else 1.0 / (Nz - 1)
        for scheme in (P.PADE6, P.PENTA10, P.CD2):
            op = LineOp.deriv(scheme, 2, bc, dims, dd)
            kname = op.kernel()
            _say(f"long z deriv shape={shape} bc={bc} scheme={scheme} kernel={kname}")
            pair = ((Nz % 32 == 0 and ((bc == P.PERIODIC and scheme != P.CD2) or Nz // 32 <= 30))
                    or (Nz % 32 != 0 and bc == P.DIRICHLET and Nz // 32 + 1 <= 29))
            assert kname.startswith("lineop_strided_pair<" if pair else "lineop_strided_multi<"), kname
            op.apply([fin], [out])
            torch.cuda.synchronize()
            assert relerr(out.cpu().numpy(), P.deriv_field(scheme, 2, bc, f, dd)) <= TOL, (scheme, bc)
        bc3 = (0, bc, bc)
        for kind, af in ((P.COMPACT8, 0.49), (P.COMPACT10, -0.49), (P.COMMN5PADE, 0.0), (P.COMMN5EXP, 0.0)):
            op = LineOp.filter(kind, af, 2, bc3, dims)
            kname = op.kernel()
            _say(f"long z filter shape={shape} bc3={bc3} kind={kind} kernel={kname}")
            pair = ((Nz % 32 == 0 and (bc == P.PERIODIC or Nz // 32 <= 30))
                    or (Nz % 32 != 0 and bc == P.DIRICHLET and Nz // 32 + 1 <= 29))
            assert kname.startswith("lineop_strided_pair<" if pair else "lineop_strided_multi<"), kname
            ref = P.filter_field(kind, af, 2, list(bc3), f)
            op.apply([fin], [out])
            torch.cuda.synchronize()
            assert relerr(out.cpu().numpy(), ref) <= TOL, (kind, bc)
            fl = torch.from_numpy(f).cuda()
            op.apply([fl], [fl])
            torch.cuda.synchronize()
            assert relerr(fl.cpu().numpy(), ref) <= TOL, (kind, bc, "in place")


def test_batched_fields_and_inplace_filter(torch_cuda):
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    shape = (24, 32, 40)
    dims = (40, 32, 24)
    fs = [field(shape, 10 + i) for i in range(5)]
    ins = [torch.from_numpy(f).cuda() for f in fs]
    outs = [torch.empty_like(t) for t in ins]
    _say("batched penta10 y-derivative, 5 fields")
    op = LineOp.deriv(P.PENTA10, 1, P.PERIODIC, dims, 1.0 / 32)
    op.apply(ins, outs)
    torch.cuda.synchronize()
    for f, o in zip(fs, outs):
        assert relerr(o.cpu().numpy(), P.deriv_field(P.PENTA10, 1, P.PERIODIC, f, 1.0 / 32)) <= TOL
    _say("in-place compact8 z-filter, 5 fields")
    fl = LineOp.filter(P.COMPACT8, 0.49, 2, (0, 0, 0), dims)
    fl.apply(ins, ins)
    torch.cuda.synchronize()
    for f, o in zip(fs, ins):
        assert relerr(o.cpu().numpy(), P.filter_field(P.COMPACT8, 0.49, 2, [0, 0, 0], f)) <= TOL


@pytest.mark.parametrize("kind,af", [(P.COMPACT8, 0.49), (P.COMMN5PADE, 0.0)])
@pytest.mark.parametrize("bc3", [(0, 0, 0), (1, 1, 1)])
def test_inplace_filter_fast_kernels(torch_cuda, kind, af, bc3):
    """In-place filtering through the fast kernels (evenly chunked lines in all
    three directions), with Dirichlet pass-through lines when bc3 asks."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    shape = (48, 32, 64)
    dims = (64, 32, 48)
    fs = [field(shape, 30 + i) for i in range(3)]
    for d in range(3):
        fl = LineOp.filter(kind, af, d, bc3, dims)
        kname = fl.kernel()
        _say(f"in-place filter kind={kind} bc3={bc3} dir={d} kernel={kname}")
        assert kname.startswith("lineop_contig" if d == 0 else "lineop_strided"), kname
        ts = [torch.from_numpy(f).cuda() for f in fs]
        fl.apply(ts, ts)
        torch.cuda.synchronize()
        for f, t in zip(fs, ts):
            assert relerr(t.cpu().numpy(), P.filter_field(kind, af, d, list(bc3), f)) <= TOL


def test_kernel_choice(torch_cuda):
    """Evenly chunked lines take the fast kernels, periodic CD2 (whose wrap
    rows flip the sign of its interior) included, and so do ragged lines of at
    least three chunks whose short chunk holds 2W rows; other ragged lines and
    TPB take the LDS-tiled kernel.  x-lines take 16-row chunks."""
    from mpi_3dcompact_amd.ops import LineOp
    dims = (64, 17, 48)
    assert LineOp.deriv(P.PENTA10, 0, P.PERIODIC, dims, 1 / 64).kernel() == "lineop_contig<2, 3, 16, true>"
    assert LineOp.deriv(P.PENTA10, 0, P.PERIODIC, (256, 17, 48), 1 / 256).kernel() == "lineop_contig<2, 3, 16, true>"
    assert LineOp.deriv(P.PADE6, 0, P.PERIODIC, (256, 17, 48), 1 / 256).kernel() == "lineop_contig<1, 2, 16, true>"
    assert LineOp.filter(P.COMMN5PADE, 0.0, 0, (0, 0, 0), (256, 17, 48)).kernel() == "lineop_contig<2, 5, 16, true>"
    assert LineOp.deriv(P.PENTA10, 1, P.PERIODIC, (17, 256, 48), 1 / 256).kernel() == "lineop_strided<2, 3, 32, true>"
    assert LineOp.deriv(P.PENTA10, 2, P.DIRICHLET, dims, 1 / 47).kernel() == "lineop_strided<2, 3, 16, false>"
    assert LineOp.deriv(P.PADE6, 1, P.PERIODIC, dims, 1 / 17).kernel().startswith("lineop_kernel<")
    assert LineOp.deriv(P.CD2, 0, P.PERIODIC, dims, 1 / 64).kernel() == "lineop_contig<0, 1, 16, true>"
    assert LineOp.deriv(P.CD2, 2, P.PERIODIC, dims, 1 / 48).kernel() == "lineop_strided<0, 1, 16, true>"
    assert LineOp.deriv(P.CD2, 1, P.PERIODIC, dims, 1 / 17).kernel().startswith("lineop_kernel<")
    assert LineOp.deriv(P.PENTA10, 0, P.PERIODIC, (150, 17, 48), 1 / 150).kernel() == \
        "lineop_contig<2, 3, 16, true, ragged>"
    assert LineOp.deriv(P.PENTA10, 2, P.DIRICHLET, (16, 17, 310), 1 / 309).kernel() == \
        "lineop_strided<2, 3, 32, false, ragged>"
    assert LineOp.deriv(P.PENTA10, 2, P.DIRICHLET, (16, 17, 131), 1 / 130).kernel().startswith("lineop_kernel<")


@pytest.mark.parametrize("scheme", [P.PADE6, P.PENTA10, P.CD2])
def test_tpb_wrap_values(torch_cuda, scheme):
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    shape = (16, 20, 24)
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    rng = np.random.default_rng(3)
    f = field(shape, 4)
    fin = torch.from_numpy(f).cuda()
    for d in range(3):
        nlines = f.size // dims[d]
        wraps = [rng.normal(size=nlines) for _ in range(6)]
        _say(f"tpb scheme={scheme} dir={d}")
        op = LineOp.deriv(scheme, d, P.PERIODIC, dims, 1.0 / dims[d])
        out = torch.empty_like(fin)
        op.apply_tpb(fin, out, [torch.from_numpy(w).cuda() for w in wraps])
        torch.cuda.synchronize()
        ref = P.deriv_field_tpb(scheme, d, f, 1.0 / dims[d], wraps)
        assert relerr(out.cpu().numpy(), ref) <= TOL


def test_kernel_level_config_256(torch_cuda):
    """256^3 curvilinear kernel-level workload: PENTA10 derivatives (periodic xi,
    Dirichlet eta/zeta) and COMPACT8 filters (alpha_f = 0.49) on 5 fields.
    Full-size properties plus oracle parity on sampled lines."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    N = 256
    dims = (N, N, N)
    bc = (P.PERIODIC, P.DIRICHLET, P.DIRICHLET)
    rng = np.random.default_rng(11)
    i = torch.arange(N, dtype=torch.float64, device="cuda")
    # xi is periodic (spacing 1/N); eta and zeta are Dirichlet (spacing 1/(N-1))
    xi = torch.arange(N, dtype=torch.float64, device="cuda") / N
    z, y, x = torch.meshgrid(i / (N - 1), i / (N - 1), xi, indexing="ij")
    f = torch.ones((N, N, N), dtype=torch.float64, device="cuda")
    for m in range(32):
        k = rng.integers(-16, 17, size=3)
        f += 1e-2 / np.sqrt(32) * torch.cos(2 * np.pi * (int(k[0]) * x + int(k[1]) * y + int(k[2]) * z)
                                            + float(rng.uniform(0, 2 * np.pi)))
    host = f.cpu().numpy()
    for d in range(3):
        dd = 1.0 / N if bc[d] == P.PERIODIC else 1.0 / (N - 1)
        _say(f"256^3 penta10 dir={d}")
        op = LineOp.deriv(P.PENTA10, d, bc[d], dims, dd)
        out = torch.empty_like(f)
        op.apply([f], [out])
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        # sampled lines against the oracle
        for (a, b) in [(0, 0), (17, 200), (255, 128), (101, 3)]:
            if d == 0:
                line, got = host[a, b, :], o[a, b, :]
            elif d == 1:
                line, got = host[a, :, b], o[a, :, b]
            else:
                line, got = host[:, a, b], o[:, a, b]
            ref = P.deriv_line(P.PENTA10, bc[d], line, dd)
            assert np.max(np.abs(got - ref)) <= TOL * max(np.max(np.abs(ref)), 1.0)
        _say(f"256^3 compact8 dir={d}")
        fl = LineOp.filter(P.COMPACT8, 0.49, d, bc, dims)
        fo = torch.empty_like(f)
        fl.apply([f], [fo])
        torch.cuda.synchronize()
        o = fo.cpu().numpy()
        ref = P.filter_field(P.COMPACT8, 0.49, d, list(bc), host[:8] if d != 2 else host)
        if d != 2:
            # the 8-plane oracle slab treats plane 7 as a Dirichlet face; compare planes 0..6
            assert relerr(o[:7], ref[:7]) <= TOL
        else:
            assert relerr(o, ref) <= TOL


# lines that split evenly into 20-, 24-, 28-row chunks (any boundary) or 12-row
# ones (periodic): the fast kernels take them (src/Pade6.cpp etc. restated by the
# oracle; the chunking is this library's)
EVEN_SHAPES = [((24, 56, 40), {0: 20, 1: 28, 2: 12}), ((36, 72, 24), {0: 12, 1: 24, 2: 12}),
               ((12, 16, 200), {0: 20, 1: 16, 2: 16})]


@pytest.mark.parametrize("shape,mm", EVEN_SHAPES, ids=["40x56x24", "24x72x36", "200x16x12"])
def test_uneven_sizes_take_the_fast_kernels(torch_cuda, shape, mm):
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 7)
    fin = torch.from_numpy(f).cuda()
    out = torch.empty_like(fin)
    for d in range(3):
        n = dims[d]
        for bc in (P.PERIODIC, P.DIRICHLET):
            for scheme in (P.PADE6, P.PENTA10):
                dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
                op = LineOp.deriv(scheme, d, bc, dims, dd)
                kname = op.kernel()
                _say(f"deriv scheme={scheme} bc={bc} shape={shape} dir={d} kernel={kname}")
                fast = mm[d] >= 16 or bc == P.PERIODIC
                if n % mm[d] == 0 and n // mm[d] >= 2 and fast:  # a periodic line needs two chunks
                    want = ("lineop_contig<" if d == 0 else "lineop_strided<")
                    assert kname.startswith(want) and f", {mm[d]}, " in kname, kname
                op.apply([fin], [out])
                torch.cuda.synchronize()
                assert relerr(out.cpu().numpy(), P.deriv_field(scheme, d, bc, f, dd)) <= TOL, (scheme, bc, d)
        for kind, af in ((P.COMPACT8, 0.49), (P.COMPACT10, 0.49), (P.COMMN5PADE, 0.0)):
            for bc3 in ((0, 0, 0), (1, 1, 1)):
                _say(f"filter kind={kind} bc3={bc3} shape={shape} dir={d}")
                op = LineOp.filter(kind, af, d, bc3, dims)
                op.apply([fin], [out])
                torch.cuda.synchronize()
                assert relerr(out.cpu().numpy(), P.filter_field(kind, af, d, list(bc3), f)) <= TOL, (kind, bc3, d)


def test_eight_row_chunks(torch_cuda):
    """88-point periodic lines split into eleven 8-row chunks: the fast kernels
    take derivatives and the Compact8 filter (half-width 4 fits 8 rows);
    Compact10 and CommN5Pade (half-width 5) take the fast kernels' ragged
    layout (five 16-row chunks and one of 8 rows), as do Dirichlet lines of
    88 points.  24-point z-lines split into 12-row chunks."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    shape = (24, 88, 88)
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    mm = {0: 8, 1: 8, 2: 12}
    f = field(shape, 11)
    fin = torch.from_numpy(f).cuda()
    out = torch.empty_like(fin)
    for d in range(3):
        fam = "lineop_contig<" if d == 0 else "lineop_strided<"
        for bc in (P.PERIODIC, P.DIRICHLET):
            n = dims[d]
            for scheme in (P.PADE6, P.PENTA10):
                dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
                op = LineOp.deriv(scheme, d, bc, dims, dd)
                kname = op.kernel()
                _say(f"deriv scheme={scheme} bc={bc} dir={d} kernel={kname}")
                if bc == P.PERIODIC:
                    assert kname.startswith(fam) and kname.endswith(f", {mm[d]}, true>"), kname
                elif mm[d] == 8:
                    assert kname.startswith(fam) and kname.endswith(", 16, false, ragged>"), kname
                op.apply([fin], [out])
                torch.cuda.synchronize()
                assert relerr(out.cpu().numpy(), P.deriv_field(scheme, d, bc, f, dd)) <= TOL, (scheme, bc, d)
        for kind, af in ((P.COMPACT8, 0.49), (P.COMPACT10, 0.49), (P.COMMN5PADE, 0.0)):
            op = LineOp.filter(kind, af, d, (0, 0, 0), dims)
            kname = op.kernel()
            _say(f"filter kind={kind} dir={d} kernel={kname}")
            if kind == P.COMPACT8:
                assert kname.startswith(fam) and kname.endswith(f", {mm[d]}, true>"), kname
            elif mm[d] == 8:
                assert kname.startswith(fam) and kname.endswith(", 16, true, ragged>"), kname
            op.apply([fin], [out])
            torch.cuda.synchronize()
            assert relerr(out.cpu().numpy(), P.filter_field(kind, af, d, [0, 0, 0], f)) <= TOL, (kind, d)


@pytest.mark.parametrize("nx,ny,nz", [(384, 20, 4), (512, 20, 4), (576, 13, 11)])
def test_long_periodic_x_lines(torch_cuda, nx, ny, nz):
    """x-lines longer than sixteen 16-row chunks: the contiguous kernel's
    workgroup is 8 lines x P chunks above 256 points (192 to 288 lanes), so
    it still takes them with 16-row chunks; 143 lines leave the last
    workgroup of 8 partly empty."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    shape = (nz, ny, nx)
    dims = (nx, ny, nz)
    f = field(shape, 12)
    fin = torch.from_numpy(f).cuda()
    out = torch.empty_like(fin)
    for scheme in (P.PADE6, P.PENTA10):
        op = LineOp.deriv(scheme, 0, P.PERIODIC, dims, 1.0 / nx)
        kname = op.kernel()
        _say(f"deriv scheme={scheme} nx={nx} kernel={kname}")
        assert kname.startswith("lineop_contig<") and kname.endswith(", 16, true>"), kname
        op.apply([fin], [out])
        torch.cuda.synchronize()
        assert relerr(out.cpu().numpy(), P.deriv_field(scheme, 0, P.PERIODIC, f, 1.0 / nx)) <= TOL, scheme
    for kind, af in ((P.COMPACT8, 0.49), (P.COMPACT10, 0.49), (P.COMMN5PADE, 0.0)):
        op = LineOp.filter(kind, af, 0, (0, 0, 0), dims)
        kname = op.kernel()
        _say(f"filter kind={kind} nx={nx} kernel={kname}")
        assert kname.startswith("lineop_contig<") and kname.endswith(", 16, true>"), kname
        op.apply([fin], [out])
        torch.cuda.synchronize()
        assert relerr(out.cpu().numpy(), P.filter_field(kind, af, 0, [0, 0, 0], f)) <= TOL, kind


def test_long_dirichlet_x_lines(torch_cuda):
    """288-point Dirichlet x-lines: 16-row chunks would need 18 x 64 lanes, so
    the plan's 32-row chunks are tried next (9 x 64 lanes) and take the
    contiguous kernel when its 64-line tile fits the workgroup's LDS."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    nx = 288
    shape = (4, 20, nx)
    dims = (nx, 20, 4)
    f = field(shape, 13)
    fin = torch.from_numpy(f).cuda()
    out = torch.empty_like(fin)
    lds = getattr(torch.cuda.get_device_properties(0), "shared_memory_per_block", "?")
    for scheme in (P.PADE6, P.PENTA10):
        dd = 1.0 / (nx - 1)
        op = LineOp.deriv(scheme, 0, P.DIRICHLET, dims, dd)
        kname = op.kernel()
        _say(f"deriv scheme={scheme} nx={nx} dirichlet kernel={kname} lds/block={lds}")
        assert kname.startswith("lineop_kernel<") or (kname.startswith("lineop_contig<") and
                                                       kname.endswith(", 32, false>")), kname
        op.apply([fin], [out])
        torch.cuda.synchronize()
        assert relerr(out.cpu().numpy(), P.deriv_field(scheme, 0, P.DIRICHLET, f, dd)) <= TOL, scheme


def test_fields_above_2gib(torch_cuda):
    """A 650^3 field (2.2 GB, above 2^31 bytes: the strided kernels' buffer
    offsets and record counts are unsigned 32-bit): Penta10 derivatives in
    every direction, periodic and Dirichlet, y and z through the fast kernels
    (ragged lines: lineop_strided_multi, lineop_strided_pair); lines at
    both ends of the field (byte offsets near 2.2e9) equal bitwise the same
    lines of small boxes cut from it, which the other tests hold to the oracle."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    n = 650
    g = torch.Generator(device="cuda").manual_seed(3)
    f = torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    out = torch.empty_like(f)
    ends = torch.tensor([0, 1, 2, 3, n - 4, n - 3, n - 2, n - 1], device="cuda")
    for bc in (P.PERIODIC, P.DIRICHLET):
        dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
        for direction in (0, 1, 2):
            op = LineOp.deriv(P.PENTA10, direction, bc, (n, n, n), dd)
            kname = op.kernel()
            _say(f"650^3 field dir={direction} bc={bc} kernel={kname}")
            if direction != 0:  # (ragged x-lines of more than 16 chunks take the LDS-tiled kernel)
                assert not kname.startswith("lineop_kernel<"), kname
            op.apply([f], [out])
            torch.cuda.synchronize()
            # the lines of direction `direction` through the end rows of the other two axes
            if direction == 0:
                sub, ref = f[ends][:, ends], out[ends][:, ends]
            elif direction == 1:
                sub, ref = f[ends][:, :, ends], out[ends][:, :, ends]
            else:
                sub, ref = f[:, ends][:, :, ends], out[:, ends][:, :, ends]
            sub = sub.contiguous()
            Nz, Ny, Nx = sub.shape
            small = LineOp.deriv(P.PENTA10, direction, bc, (Nx, Ny, Nz), dd)
            assert small.kernel() == kname, (small.kernel(), kname)
            o2 = torch.empty_like(sub)
            small.apply([sub], [o2])
            torch.cuda.synchronize()
            assert torch.equal(o2, ref.contiguous()), (direction, bc)
            del op, small
    del f, out
    torch.cuda.empty_cache()
