"""Image outputs on the CPU: the oracle's colour maps and pixel grid pinned by
what they must give (the rainbow's five nodes, greyscale, the blue-white-red
ends and middle, the equal-area pixel centres), then the library's host-side
renderer and PNG writer (mc3d_image_write) against the oracle, bit for bit,
through an independent PNG decoder that checks every CRC."""
import numpy as np
import pytest

from oracle import orc_images as I


def test_rainbow_nodes_and_greyscale():
    w = 0.2500001
    nodes = [(0, 0, 255), (0, 255, 255), (0, 255, 0), (255, 255, 0)]
    for k, rgb in enumerate(nodes):
        assert I.colormap(I.RAINBOW, k * w) == rgb
    # the top end sits just short of the last node: red, green rounding down from 0
    r, g, b = I.colormap(I.RAINBOW, 1.0)
    assert r >= 254 and g <= 1 and b == 0
    # half-way between blue and cyan
    assert I.colormap(I.RAINBOW, 0.5 * w) == (0, 127, 255)
    for f, v in ((0.0, 0), (0.5, 127), (1.0, 255)):
        assert I.colormap(I.GREYSCALE, f) == (v, v, v)


def test_blue_white_red_ends():
    assert I.colormap(I.BWR, 0.0) == (19, 119, 255)
    r, g, b = I.colormap(I.BWR, 0.5)
    assert min(r, g, b) >= 238          # near white in the middle
    r, g, b = I.colormap(I.BWR, 1.0)
    assert r > 200 and g < 70 and b < 60  # red end
    reds = [I.colormap(I.BWR, f)[0] for f in np.linspace(0, 1, 101)]
    blues = [I.colormap(I.BWR, f)[2] for f in np.linspace(0, 1, 101)]
    assert all(np.diff(reds[:50]) >= 0) and all(np.diff(blues[55:]) <= 0)


def test_pixel_centres():
    # d = 1 when n - 1 = size + 1: pixel p samples index p + 1 (centre 0.5 + p, rounded up)
    src = I.pixel_sources(9, 9, 7, 7)
    assert src[0, 0] == 1 * 9 + 1 and src[6, 6] == 7 * 9 + 7
    # a wide plane in a square picture: the larger spacing wins, the top rows fall off the plane
    src = I.pixel_sources(33, 5, 15, 15)
    d = max(32 / 16, 4 / 16)
    assert d == 2.0
    # centres d/2 + d p = 1 + 2p: row 0 samples i2 = 1, pixel 3 samples i1 = 7
    assert src[0, 3] == 1 * 33 + 7
    # rows 2 and up reach i2 = 5 > 4: background
    assert np.all(src[2:] == -1) and np.all(src[:2] >= 0)


@pytest.mark.parametrize("cm", [I.GREYSCALE, I.RAINBOW, I.BWR])
@pytest.mark.parametrize("vrange", [None, (-0.5, 0.8)])
def test_library_png_matches_the_oracle(tmp_path, cm, vrange):
    from mpi_3dcompact_amd.solver import image_write
    rng = np.random.default_rng(cm + 7)
    n2, n1 = 19, 27
    a1, a2 = np.meshgrid(np.linspace(0, 1, n1), np.linspace(0, 1, n2))
    plane = np.sin(3 * a1) * np.cos(2 * a2) + 0.05 * rng.standard_normal((n2, n1))
    path = tmp_path / "P.000000.png"
    image_write(path, plane, 40, 23, vrange=vrange, colormap=cm)
    got = I.read_png(path)
    want = I.render(plane, 40, 23, vrange, cm)
    assert got.shape == (23, 40, 3)
    assert np.array_equal(got, want)
    assert np.any(np.all(got == I.BACKGROUND, axis=2)) == np.any(I.pixel_sources(n1, n2, 40, 23) < 0)


def test_flat_plane_and_bad_arguments(tmp_path):
    from mpi_3dcompact_amd._lib import Mc3dError
    from mpi_3dcompact_amd.solver import image_write
    path = tmp_path / "F.png"
    image_write(path, np.full((6, 6), 3.0), 8, 8, colormap="RAINBOW")
    assert np.array_equal(I.read_png(path), I.render(np.full((6, 6), 3.0), 8, 8, None, I.RAINBOW))
    with pytest.raises(Mc3dError):
        image_write(path, np.zeros((4, 4)), 0, 8)
    with pytest.raises(Mc3dError):
        image_write(path, np.zeros((4, 4)), 8, 8, colormap=5)
    with pytest.raises(Mc3dError):
        image_write(tmp_path / "missing" / "x.png", np.zeros((4, 4)), 8, 8)
