"""The cylinder O-grid of src/AlgebraicSingleBlockMesh.hpp:210-263 (CPU): the
radial table (wall at r = 0.5, first offset 8e-4, growth 1.04 for the first 50
layers then 1.02), the tanh clustering of theta over [0, 2 pi), z = pi k / Nz,
and the uniform-inflow state of src/MPI_3DCompact.cpp:200-206."""
import numpy as np
import pytest

from mpi_3dcompact_amd import cylinder


def test_radial_table():
    r = cylinder.radii(60)
    dr = np.diff(r)
    assert r[0] == 0.5 and dr[0] == pytest.approx(8e-4, rel=1e-12)
    np.testing.assert_allclose(dr[1:49] / dr[:48], 1.04, rtol=1e-9)
    np.testing.assert_allclose(dr[49:] / dr[48:-1], 1.02, rtol=1e-9)
    # closed form of the first 50 layers: a geometric series
    assert r[49] == pytest.approx(0.5 + 8e-4 * (1.04 ** 49 - 1) / 0.04, rel=1e-13)
    with pytest.raises(ValueError):
        cylinder.radii(49)


def test_o_grid_coordinates():
    Nx, Ny, Nz = 24, 52, 12
    x, y, z = cylinder.cylinder_mesh(Nx, Ny, Nz)
    assert x.shape == (Nz, Ny, Nx)
    r = np.hypot(x, y)
    np.testing.assert_allclose(r, np.broadcast_to(cylinder.radii(Ny)[None, :, None], r.shape), rtol=1e-14)
    theta = np.mod(np.arctan2(y[0, 0], x[0, 0]), 2 * np.pi)
    assert theta[0] == 0.0 and np.all(np.diff(theta) > 0) and theta[-1] < 2 * np.pi
    # tanh clustering: symmetric about the half-way point, finer at both ends (the wrap)
    np.testing.assert_allclose(theta[Nx // 2], np.pi, atol=1e-14)
    d = np.diff(np.append(theta, 2 * np.pi))
    assert d[0] < d[Nx // 2 - 1] and d[-1] < d[Nx // 2]
    np.testing.assert_allclose(d, d[::-1], rtol=1e-12)
    np.testing.assert_allclose(z[:, 0, 0], np.pi * np.arange(Nz) / Nz, rtol=0, atol=1e-15)


def test_inflow_state():
    ic = cylinder.initial_condition((2, 3, 4))
    assert np.all(ic["U"] == 0.2) and np.all(ic["rho"] == 1.0) and np.all(ic["p"] == 1.0 / 1.4)
