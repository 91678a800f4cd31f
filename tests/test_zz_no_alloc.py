"""mc3d_cfl_dt_noalloc (the CFL time step with its partial maxima in the
library's static scratch array) against mc3d_cfl_dt (which allocates them),
through the C ABI: bitwise equal dt and maximum speed on two calls in a row,
the second on fewer and smaller boxes with lower wave speeds, so partial
maxima left in the scratch array by the first call would show; and the box
limit is refused with MC3D_EINVAL.
"""
import sys

import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch


def _fields(torch, n, members, speed, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    tot = n * members

    def uni(lo, hi):
        return (lo + (hi - lo) * torch.rand(tot, generator=g, dtype=torch.float64)).cuda()

    return [uni(-speed, speed), uni(-speed, speed), uni(-speed, speed), uni(0.5, 1.5),
            uni(0.5, 1.5), uni(0.5, 1.5), uni(0.5, 1.5)]


def test_noalloc_matches_allocating_entry_twice(torch_cuda):
    from mpi_3dcompact_amd.sgs import cfl_dt
    torch = torch_cuda
    spacing = (1.0 / 32, 1.0 / 40, 1.0 / 48)
    # first call: 3 boxes of 32x40x48 points; second: 2 boxes of 12x16x20 with
    # wave speeds about ten times lower (stale partials would raise its maxima)
    for call, (n, members, speed, seed) in enumerate(((32 * 40 * 48, 3, 10.0, 11), (12 * 16 * 20, 2, 1.0, 12))):
        f = _fields(torch, n, members, speed, seed)
        _say(f"call {call}: {members} boxes of {n} points, speed {speed}")
        want = cfl_dt(*f, spacing, 0.5, members=members)
        out = torch.full((members, 2), float("nan"), dtype=torch.float64, device="cuda")
        got = cfl_dt(*f, spacing, 0.5, members=members, out=out, noalloc=True)
        torch.cuda.synchronize()
        w, g = want.cpu(), got.cpu()
        _say(f"  allocating {w.tolist()}\n  noalloc    {g.tolist()}")
        assert torch.equal(w, g), (call, w, g)
        # the maximum itself, from torch on the same fields, per box
        U, V, W, c, J11, J22, J33 = (t.view(members, n) for t in f)
        ms = (J11 * (U.abs() + c) / spacing[0] + J22 * (V.abs() + c) / spacing[1]
              + J33 * (W.abs() + c) / spacing[2]).max(dim=1).values.cpu()
        assert torch.allclose(g[:, 1], ms, rtol=1e-15, atol=0), (g[:, 1], ms)


def test_noalloc_refuses_more_boxes_than_its_scratch(torch_cuda):
    from mpi_3dcompact_amd._lib import Mc3dError
    from mpi_3dcompact_amd.sgs import cfl_dt
    torch = torch_cuda
    # 65 boxes of 2048 * 256 points would need 65 * 2048 partials: refused
    # before anything is queued
    n = 2048 * 256
    f = [torch.ones(n * 65, dtype=torch.float64, device="cuda") for _ in range(7)]
    with pytest.raises(Mc3dError, match="at most 64 boxes"):
        cfl_dt(*f, (1.0, 1.0, 1.0), 0.5, members=65, noalloc=True)
    # the allocating entry point takes them
    out = cfl_dt(*f, (1.0, 1.0, 1.0), 0.5, members=65, noalloc=False)
    torch.cuda.synchronize()
    assert torch.all(out[:, 1] == 6.0)
