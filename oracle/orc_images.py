"""Test infrastructure, not product code: a CPU restatement of the image
outputs (the checker of the library's PNG writer), and a PNG decoder.

colormap(): src/PngWriter.hpp:270-313 (RAINBOW, BWR) and the GREYSCALE case
of src/CurvilinearCSolver_Core.cpp:1729-1733; the colour scaling of
src/CurvilinearCSolver_Core.cpp:1686-1727.  The pixel placement restates the
reference's equal-area pixel grid (src/CurvilinearCSolver_Core.cpp:1494-1650:
dx = max(L1 / (N1 + 1), L2 / (N2 + 1)), centres at dx / 2 + dx * ip, row
jp = 0 at the bottom) in index space, a pixel taking the nearest grid index
and the background colour (73, 175, 205) past the plane."""
import math
import struct
import zlib

import numpy as np

GREYSCALE, RAINBOW, BWR = 0, 1, 2
BACKGROUND = (73, 175, 205)

_RAINBOW = ((0, 0, 0, 255, 255), (0, 255, 255, 255, 0), (255, 255, 0, 0, 0))
# ParaView's blue-white-red table as the reference lists it (33 steps, [0, 1])
_BWR = (
    (0.07514311, 0.247872569, 0.339526309, 0.409505078, 0.468487184, 0.520796675, 0.568724526, 0.613686735,
     0.656658579, 0.698372844, 0.739424025, 0.780330104, 0.821573924, 0.863634967, 0.907017747, 0.936129275,
     0.943467973, 0.990146732, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0.961891484, 0.916482116),
    (0.468049805, 0.498782363, 0.528909511, 0.558608486, 0.588057293, 0.617435078, 0.646924167, 0.676713218,
     0.707001303, 0.738002964, 0.769954435, 0.803121429, 0.837809045, 0.874374691, 0.913245283, 0.938743558,
     0.943498599, 0.928791426, 0.88332677, 0.833985467, 0.788626485, 0.746206642, 0.70590052, 0.667019783,
     0.6289553, 0.591130233, 0.552955184, 0.513776083, 0.472800903, 0.428977855, 0.380759558, 0.313155629,
     0.236630659),
    (1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0.983038586, 0.943398095, 0.917447482, 0.861943246,
     0.803839606, 0.750707739, 0.701389973, 0.654994046, 0.610806959, 0.568237474, 0.526775617, 0.485962266,
     0.445364274, 0.404551679, 0.363073592, 0.320428137, 0.265499262, 0.209939162),
)


def _interp(maps, w, f):
    c = int(math.floor(f / w))
    t = (f - w * c) / w
    return tuple(int((1.0 - t) * m[c] + t * m[c + 1]) for m in maps)


def colormap(cm, f):
    """(r, g, b) of f in [0, 1]."""
    if cm == RAINBOW:
        return _interp(_RAINBOW, 0.2500001, f)
    if cm == BWR:
        return _interp(tuple(tuple(int(255.0 * v) for v in m) for m in _BWR), 0.031250001, f)
    v = int(f * 255.0)
    return (v, v, v)


def pixel_sources(n1, n2, width, height):
    """(height, width) array of flat plane indices i2 * n1 + i1 (-1: background),
    row 0 = jp 0 (the bottom of the picture)."""
    d = max((n1 - 1) / (width + 1.0), (n2 - 1) / (height + 1.0))
    src = np.full((height, width), -1, dtype=np.int64)
    for jp in range(height):
        i2 = int(math.floor(d / 2 + d * jp + 0.5))
        for ip in range(width):
            i1 = int(math.floor(d / 2 + d * ip + 0.5))
            if i1 <= n1 - 1 and i2 <= n2 - 1:
                src[jp, ip] = i2 * n1 + i1
    return src


def render(plane, width, height, vrange=None, cm=RAINBOW):
    """(height, width, 3) uint8 image, top row first, of an (n2, n1) plane."""
    n2, n1 = plane.shape
    flat = np.ascontiguousarray(plane, dtype=np.float64).ravel()
    src = pixel_sources(n1, n2, width, height)
    if vrange is not None:
        lo, hi = vrange
    else:
        lo, hi = 100000000.0, -100000000.0
        for s in src[src >= 0]:
            lo, hi = min(lo, flat[s]), max(hi, flat[s])
    img = np.empty((height, width, 3), dtype=np.uint8)
    for jp in range(height):
        for ip in range(width):
            s = src[jp, ip]
            if s < 0:
                rgb = BACKGROUND
            else:
                f = (flat[s] - lo) / (hi - lo) if hi != lo else 0.0
                if hi - lo < 1e-6:
                    f = 0.0
                rgb = colormap(cm, min(1.0, max(0.0, f)))
            img[height - 1 - jp, ip] = rgb
    return img


def read_png(path):
    """8-bit RGB, non-interlaced PNG -> (height, width, 3) uint8, checking every
    chunk's CRC; rows with filter type 0 only (what the writer emits)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG signature"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(typ + body) & 0xffffffff == crc, f"bad CRC in {typ!r}"
        if typ == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
        pos += 12 + n
    w, h, depth, ctype, comp, filt, inter = hdr
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(idat)
    assert len(raw) == h * (3 * w + 1)
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 3 * w + 1)
    assert np.all(rows[:, 0] == 0), "unexpected row filter"
    return rows[:, 1:].reshape(h, w, 3).copy()
