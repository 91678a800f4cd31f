"""Slices and image outputs of a running solver on the GPU: the device slice
equals the field's index plane bit for bit in every direction, and the PNG
files land at the reference's steps (the first step, then every interval,
src/CurvilinearCSolver_Core.cpp:1442-1492) with the pixels the oracle draws
from the field at that step."""
import sys

import numpy as np
import pytest

from mpi_3dcompact_amd import hit
from oracle import orc_images as I

pytestmark = pytest.mark.gpu

CASE = dict(fd="PADE6", filt="COMPACT10", rk="TVDRK3", mu_ref=1e-3)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def make(dims, **kw):
    from mpi_3dcompact_amd.solver import Solver
    Nx, Ny, Nz = dims
    g = Solver(hit.options_text(Nx, dims=dims, **CASE, **kw), (Nz, Ny, Nx))
    d = 1 / (2 * np.pi)   # x = 2 pi xi in every direction
    g.set_metrics([np.full((Nz, Ny, Nx), m) for m in (d ** 3, d, 0.0, 0.0, 0.0, d, 0.0, 0.0, 0.0, d)])
    rng = np.random.default_rng(4)
    shape = (Nz, Ny, Nx)
    g.set_initial(dict(rho=1 + 0.01 * rng.standard_normal(shape), U=0.1 * rng.standard_normal(shape),
                       V=0.1 * rng.standard_normal(shape), W=0.1 * rng.standard_normal(shape),
                       p=np.full(shape, 1 / 1.4)))
    return g


def test_slices_are_index_planes(gpu):
    dims = (20, 16, 12)
    _say("case: slices of a 20x16x12 field")
    g = make(dims)
    g.step(1)
    for name in ("rho", "U", "p"):
        f = g.field(name)
        for k in (0, 5, 11):
            assert np.array_equal(g.slice(name, 2, k), f[k]), (name, 2, k)
        for j in (0, 7, 15):
            assert np.array_equal(g.slice(name, 1, j), f[:, j, :].T), (name, 1, j)
        for i in (0, 9, 19):
            assert np.array_equal(g.slice(name, 0, i), f[:, :, i]), (name, 0, i)
    from mpi_3dcompact_amd._lib import Mc3dError
    with pytest.raises(Mc3dError):
        g.slice("rho", 2, 12)
    with pytest.raises(Mc3dError):
        g.slice("nothing", 2, 0)


def test_image_outputs_follow_the_steps(gpu, tmp_path):
    dims = (24, 24, 16)
    _say("case: image outputs on 24x24x16, 4 steps")
    g = make(dims, max_steps=4)
    g.set_output(str(tmp_path))
    g.add_image("U", "U", plane=2, fraction=0.5, width=30, height=30, interval=2, vrange=(-0.2, 0.2),
                colormap="BWR")
    g.add_image("rho", "RHO0", plane=0, fraction=0.25, width=20, height=12, interval=0, colormap="GREYSCALE")
    rho0 = g.field("rho")
    g.run()
    names = sorted(p.name for p in tmp_path.iterdir() if p.suffix == ".png")
    assert names == ["RHO0.000000.png", "U.000000.png", "U.000002.png", "U.000004.png"]
    # the last U image is the current field's plane k = round(0.5 * 15) = 8
    U = g.field("U")
    assert np.array_equal(I.read_png(tmp_path / "U.000004.png"), I.render(U[8], 30, 30, (-0.2, 0.2), I.BWR))
    # the initial density, plane i = round(0.25 * 23) = 6, axes (j, k), own range
    assert np.array_equal(I.read_png(tmp_path / "RHO0.000000.png"),
                          I.render(rho0[:, :, 6], 20, 12, None, I.GREYSCALE))


def test_image_arguments_are_checked(gpu):
    from mpi_3dcompact_amd._lib import Mc3dError
    g = make((16, 16, 16))
    for bad in (dict(plane=3), dict(fraction=1.5), dict(width=0), dict(colormap=7), dict(interval=-1)):
        with pytest.raises(Mc3dError):
            g.add_image("U", "U", **bad)
    with pytest.raises(Mc3dError):
        g.add_image("nothing", "X")
