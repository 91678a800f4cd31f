"""The operators on the reference's own pencil layouts, through the C ABI.

The reference calls calc1stDerivField / filterField of direction x, y, z with
the field in the X, Y, Z pencil (src/Pade6.cpp:6-43, src/Macros.hpp:16-18):
    X pencil  index = (k*Ny + j)*Nx + i
    Y pencil  index = (i*Nz + k)*Ny + j
    Z pencil  index = (j*Nx + i)*Nz + k
and orders the wrap arrays of the transformed-periodic derivative by line as
k*Ny + j (x), i*Nz + k (y), j*Nx + i (z) (src/Pade6.cpp:56,76,95).  Here the
oracle (which works on x-fastest fields) is applied to the x-fastest field and
the result permuted to the pencil.  Tolerance: normwise relative error
<= 1e-12 per application, as for the x-fastest operators.
"""
import itertools
import sys

import numpy as np
import pytest

from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

TOL = 1e-12
XP, YP, ZP = 0, 1, 2


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


def to_layout(f, layout):
    """x-fastest (Nz, Ny, Nx) array -> the layout's array, slowest axis first."""
    if layout == XP:
        return np.ascontiguousarray(f)
    if layout == YP:   # [i, k, j]
        return np.ascontiguousarray(f.transpose(2, 0, 1))
    return np.ascontiguousarray(f.transpose(1, 2, 0))  # [j, i, k]


def field(shape, seed):
    rng = np.random.default_rng(seed)
    Nz, Ny, Nx = shape
    z, y, x = np.meshgrid(np.arange(Nz) / Nz, np.arange(Ny) / Ny, np.arange(Nx) / Nx, indexing="ij")
    f = 1.0 + 0.01 * rng.normal(size=shape)
    for _ in range(6):
        k = rng.integers(-4, 5, size=3)
        f += 0.3 * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 6.28))
    return f


# (Nz, Ny, Nx): evenly chunked lines (fast kernels) and ragged ones (LDS-tiled)
SHAPES = [(32, 48, 64), (20, 17, 24)]


def test_layout_argument_is_checked(torch_cuda):
    from mpi_3dcompact_amd._lib import Mc3dError
    from mpi_3dcompact_amd.ops import LineOp
    with pytest.raises(Mc3dError, match="layout"):
        LineOp.deriv(P.PADE6, 1, P.PERIODIC, (16, 16, 16), 1 / 16, layout=3)
    with pytest.raises(Mc3dError, match="layout"):
        LineOp.filter(P.COMPACT8, 0.49, 0, (0, 0, 0), (16, 16, 16), layout=-1)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("scheme,bc", list(itertools.product([P.PADE6, P.PENTA10, P.CD2],
                                                               [P.PERIODIC, P.DIRICHLET])))
def test_deriv_in_every_pencil(torch_cuda, scheme, bc, shape):
    """Every direction in every layout; the reference's own pairing is the
    diagonal (direction d in pencil d), which runs the contiguous kernel."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 1)
    for layout in (XP, YP, ZP):
        fin = torch.from_numpy(to_layout(f, layout)).cuda()
        for d in range(3):
            n = dims[d]
            dd = 1.0 / n if bc == P.PERIODIC else 1.0 / (n - 1)
            op = LineOp.deriv(scheme, d, bc, dims, dd, layout=layout)
            _say(f"deriv scheme={scheme} bc={bc} shape={shape} layout={layout} dir={d} kernel={op.kernel()}")
            if layout == d and n % 16 == 0 and scheme != P.CD2:
                assert op.kernel().startswith("lineop_contig<"), op.kernel()
            out = torch.empty_like(fin)
            op.apply([fin], [out])
            torch.cuda.synchronize()
            ref = to_layout(P.deriv_field(scheme, d, bc, f, dd), layout)
            assert relerr(out.cpu().numpy(), ref) <= TOL, (scheme, bc, layout, d)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind,af", [(P.COMPACT8, 0.49), (P.COMPACT10, 0.49), (P.COMPACT10, -0.49),
                                     (P.COMMN5PADE, 0.0), (P.COMMN5EXP, 0.0)])
@pytest.mark.parametrize("bc3", [(0, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0)])
def test_filter_in_every_pencil(torch_cuda, kind, af, bc3, shape):
    """Pass-through lines (Dirichlet faces of the other directions,
    src/Compact8Filter.cpp:148-222) follow the layout's axes."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    f = field(shape, 2)
    for layout in (XP, YP, ZP):
        fin = torch.from_numpy(to_layout(f, layout)).cuda()
        for d in range(3):
            _say(f"filter kind={kind} af={af} bc3={bc3} shape={shape} layout={layout} dir={d}")
            op = LineOp.filter(kind, af, d, bc3, dims, layout=layout)
            out = torch.empty_like(fin)
            op.apply([fin], [out])
            torch.cuda.synchronize()
            ref = to_layout(P.filter_field(kind, af, d, list(bc3), f), layout)
            assert relerr(out.cpu().numpy(), ref) <= TOL, (kind, bc3, layout, d)
            # in place, as doAveraging's filters may be called
            tin = fin.clone()
            op.apply([tin], [tin])
            torch.cuda.synchronize()
            assert relerr(tin.cpu().numpy(), ref) <= TOL, (kind, bc3, layout, d, "in place")


@pytest.mark.parametrize("scheme", [P.PADE6, P.PENTA10])
def test_tpb_in_reference_order(torch_cuda, scheme):
    """calc1stDerivField_TPB with the field in the direction's pencil and the
    wrap arrays in the reference's line order: x k*Ny + j, y i*Nz + k,
    z j*Nx + i."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    shape = (16, 20, 24)
    Nz, Ny, Nx = shape
    dims = (Nx, Ny, Nz)
    rng = np.random.default_rng(3)
    f = field(shape, 4)
    for d in range(3):
        nlines = f.size // dims[d]
        ref_wraps = [rng.normal(size=nlines) for _ in range(6)]
        # the oracle numbers y-lines k*Nx + i; the reference i*Nz + k
        if d == 1:
            orc_wraps = [w.reshape(Nx, Nz).T.reshape(-1).copy() for w in ref_wraps]
        else:
            orc_wraps = ref_wraps
        _say(f"tpb scheme={scheme} dir={d} in its pencil")
        op = LineOp.deriv(scheme, d, P.PERIODIC, dims, 1.0 / dims[d], layout=d)
        fin = torch.from_numpy(to_layout(f, d)).cuda()
        out = torch.empty_like(fin)
        op.apply_tpb(fin, out, [torch.from_numpy(w).cuda() for w in ref_wraps])
        torch.cuda.synchronize()
        ref = to_layout(P.deriv_field_tpb(scheme, d, f, 1.0 / dims[d], orc_wraps), d)
        assert relerr(out.cpu().numpy(), ref) <= TOL, d


def test_pencil_256_penta10_compact8(torch_cuda):
    """The kernel-level workload's operators (PENTA10 periodic xi / Dirichlet
    eta, zeta; COMPACT8 alpha_f 0.49) on 256^3 fields in each direction's own
    pencil, against the x-fastest operators of this library (which the other
    tests pin to the oracle): equal to the last bit is not required (the same
    partitioned solve, possibly other chunk sizes), 1e-12 normwise is."""
    from mpi_3dcompact_amd.ops import LineOp
    torch = torch_cuda
    N = 256
    dims = (N, N, N)
    bc = (P.PERIODIC, P.DIRICHLET, P.DIRICHLET)
    g = torch.Generator(device="cuda").manual_seed(5)
    f = 1.0 + 0.01 * torch.rand((N, N, N), dtype=torch.float64, device="cuda", generator=g)
    perm = {0: (0, 1, 2), 1: (2, 0, 1), 2: (1, 2, 0)}
    for d in range(3):
        dd = 1.0 / N if bc[d] == P.PERIODIC else 1.0 / (N - 1)
        ref_op = LineOp.deriv(P.PENTA10, d, bc[d], dims, dd)
        ref = torch.empty_like(f)
        ref_op.apply([f], [ref])
        pen = f.permute(*perm[d]).contiguous()
        op = LineOp.deriv(P.PENTA10, d, bc[d], dims, dd, layout=d)
        _say(f"256^3 penta10 dir={d} pencil kernel={op.kernel()}")
        out = torch.empty_like(pen)
        op.apply([pen], [out])
        torch.cuda.synchronize()
        want = ref.permute(*perm[d]).contiguous()
        err = (out - want).abs().max().item() / want.abs().max().item()
        assert err <= TOL, (d, err)
        fl_ref = LineOp.filter(P.COMPACT8, 0.49, d, bc, dims)
        fl = LineOp.filter(P.COMPACT8, 0.49, d, bc, dims, layout=d)
        fl_ref.apply([f], [ref])
        fl.apply([pen], [out])
        torch.cuda.synchronize()
        want = ref.permute(*perm[d]).contiguous()
        err = (out - want).abs().max().item() / want.abs().max().item()
        assert err <= TOL, (d, "filter", err)
