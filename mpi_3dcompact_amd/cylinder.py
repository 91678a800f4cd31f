"""The reference driver's cylinder case: an O-grid around a cylinder of radius
0.5 (AlgebraicSingleBlockMesh::generateCylinderGrid,
src/AlgebraicSingleBlockMesh.hpp:210-263; getMesh :71-176), uniform inflow
U = 0.2 (src/MPI_3DCompact.cpp:179-205).  x (i) runs around the cylinder
(internally periodic), y (j) outwards from the wall, z (k) spanwise over pi."""
import numpy as np

GAMMA = 1.4


def radii(Ny):
    """Radial layers: 0.5, then a first offset of 0.0008 growing by 1.04 for 50
    layers and by 1.02 beyond (the reference's table needs Ny >= 50)."""
    if Ny < 50:
        raise ValueError("the cylinder grid has 50 inner layers: Ny must be at least 50")
    r = np.empty(Ny)
    r[0] = 0.5
    r[1] = r[0] + 0.0008
    for j in range(2, 50):
        r[j] = r[j - 1] + 1.04 * (r[j - 1] - r[j - 2])
    for j in range(50, Ny):
        r[j] = r[j - 1] + 1.02 * (r[j - 1] - r[j - 2])
    return r


def cylinder_mesh(Nx, Ny, Nz):
    """x, y, z of the O-grid as (Nz, Ny, Nx) arrays: theta = 2 pi s(i / Nx) with the
    tanh clustering s(b) = (1 + tanh(2 (b - 1/2)) / tanh(1)) / 2, r = radii(Ny)[j],
    z = pi k / Nz."""
    r = radii(Ny)
    b = np.arange(Nx) / Nx
    theta = 2.0 * np.pi * (0.5 * (1.0 + np.tanh(2.0 * (b - 0.5)) / np.tanh(1.0)))
    zc = np.arange(Nz) / Nz * np.pi
    x = np.broadcast_to(r[None, :, None] * np.cos(theta)[None, None, :], (Nz, Ny, Nx)).copy()
    y = np.broadcast_to(r[None, :, None] * np.sin(theta)[None, None, :], (Nz, Ny, Nx)).copy()
    z = np.broadcast_to(zc[:, None, None], (Nz, Ny, Nx)).copy()
    return x, y, z


def initial_condition(shape, U=0.2):
    """rho = 1, p = 1 / gamma, uniform U along x."""
    one = np.ones(shape)
    return dict(rho=one.copy(), U=U * one, V=0.0 * one, W=0.0 * one, p=one / GAMMA)


FACES = {"X0": "INTERNALLY_PERIODIC", "X1": "INTERNALLY_PERIODIC", "Y0": "ADIABATIC_WALL", "Y1": "SPONGE"}
DISP = [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, np.pi]]  # z periodic over pi
