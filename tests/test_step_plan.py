"""The tile plan of the one-launch step, checked on the CPU (no device): the
library's planner (sfd_plan_step, the function build_step uses) is run on the
1024^3 benchmark geometry and on ragged grids, and its tiles must

* own every cell of the thread domain [0, nx] x [0, ny] x [0, nz] exactly once;
* fit their blocks: at most 63 owned columns (one cell per lane) or 252 in
  four-cell tiles, which start on a 16-byte boundary, own a multiple of four
  columns and keep their halo column inside every component's x sweep; at
  most rows - 1 owned rows; half-width tiles (32 lanes per row, two rows per
  wave) at most 31 columns and 2 waves - 1 rows; tiles with sources at most
  kc planes (their face masks in LDS span kc + 1 planes), every tile at most
  64;
* have every E neighbour of an owned H cell computed in the block (lane + 1,
  row + 1 and plane k1 stay inside the block and the march);
* address only memory inside the padded arrays (the four-cell kernel's
  16-byte loads of the lanes past the grid included);
* carry the code for everything their cells, halo included, meet: a CPML slab
  (bit 1), a pole-state box (bit 2), a TF/SF face or nodal box (bit 4); the
  four-cell class meets none of them.
"""
import ctypes as C

import numpy as np
import pytest

from semba_fdtd_amd import _lib

KWAVE = 64


def plan(n, L, R, kc, boxes, rows=None, vec=1, vrows=8, hrows=None):
    """Tiles [cls, x0, wx, y0, wy, k0, k1] of the planner; rows: rows per
    block of classes 0-7 (default R), hrows: waves per block of their
    half-width variants (default R, the CPML + poles + sources one 8 when R
    is 16, as the library does)."""
    lib = C.CDLL(_lib.LIB_PATH)
    f = lib.sfd_plan_step
    f.restype = C.c_int
    i32p = C.POINTER(C.c_int32)
    f.argtypes = [i32p, i32p, C.c_int32, i32p, i32p, C.c_int64, C.POINTER(C.c_int64)]
    geom = np.array(list(n) + [L[0][0], L[0][1], L[1][0], L[1][1], L[2][0], L[2][1], R, kc, vec,
                               min(vrows, R)], np.int32)
    rows = list(rows or [R] * 8)
    if hrows is None:
        hrows = [R] * 7 + [8 if R > 8 else R]
    rows = np.array(rows + list(hrows), np.int32)
    bx = np.array(boxes, np.int32).reshape(-1, 7) if boxes else np.zeros((0, 7), np.int32)
    cnt = C.c_int64()
    p = lambda a: a.ctypes.data_as(i32p)
    assert f(p(geom), p(rows), len(bx), p(bx), None, 0, C.byref(cnt)) == 0
    out = np.zeros((cnt.value, 7), np.int32)
    assert f(p(geom), p(rows), len(bx), p(bx), p(out), cnt.value, C.byref(cnt)) == 0
    return out, rows


def needs(n, L, boxes, lo, hi):
    """Class bits a tile whose cells (halo included) span [lo, hi] needs: it
    reaches into a CPML slab (the E-node or H-cell layers of either side), or
    intersects a pole box / source box."""
    bits = 0
    for a in range(3):
        if (L[a][0] > 0 and lo[a] < L[a][0]) or (L[a][1] > 0 and hi[a] >= n[a] - L[a][1]):
            bits |= 1
    for b in boxes:
        if all(lo[a] <= b[3 + a] and b[a] <= hi[a] for a in range(3)):
            bits |= b[6]
    return bits


def check(n, L, R, kc, boxes, rows=None, vec=1, vrows=8):
    tiles, rows = plan(n, L, R, kc, boxes, rows, vec, vrows)
    nx, ny, nz = n
    # row layout of fdtd_create: 32-element lead-in, 128-byte pitch from 256
    # cells wide; below that a 4-element lead-in and a 16-byte pitch
    xoff, px = (32, (32 + nx + 8 + 31) // 32 * 32) if nx >= 256 else (4, (4 + nx + 8 + 3) // 4 * 4)
    own = np.zeros((nz + 1, ny + 1, nx + 1), np.uint8)
    yee = 0
    for cls, x0, wx, y0, wy, k0, k1 in tiles:
        assert wx >= 1 and wy >= 1 and k1 > k0 >= 0 and k1 <= nz + 1
        assert x0 + wx <= nx + 1 and y0 + wy <= ny + 1
        assert k1 - k0 <= 64
        if cls & 4 and cls < 8:
            assert k1 - k0 <= kc
        if cls >= 9:
            assert wx <= KWAVE // 2 - 1              # lane wx of the 32 is the halo lane
            assert wy <= 2 * rows[cls - 1] - 1       # two rows per wave; row wy is the halo
        elif cls == 8:
            e = min(vrows, R)
            assert x0 % 4 == 0 and wx % 4 == 0 and wx <= 4 * (KWAVE - 1)
            assert x0 + wx + 1 <= nx               # halo column inside every x sweep
            assert wy <= e - 1
            # lane l loads 16 bytes at min(x0 + 4l, (nx + 3) & ~3): inside the row
            ic = np.minimum(x0 + 4 * np.arange(KWAVE), (nx + 3) & ~3)
            assert (ic % 4 == 0).all() and (xoff + ic + 3 < px).all()
            assert wx // 4 < KWAVE                   # the halo lane exists
        else:
            assert wx <= KWAVE - 1                   # lane wx is the halo lane
            assert wy <= rows[cls] - 1               # row wy is the halo row
        lo = (x0, y0, k0)
        hi = (x0 + wx, y0 + wy, k1)
        need = needs(n, L, boxes, lo, hi)
        if cls == 8:
            assert need == 0, (cls, lo, hi, need)
        else:
            bits = cls - 9 if cls >= 9 else cls
            assert need & ~bits == 0, (cls, lo, hi, need)
        own[k0:k1, y0:y0 + wy, x0:x0 + wx] += 1
        yee += max(0, min(x0 + wx, nx) - x0) * max(0, min(y0 + wy, ny) - y0) * max(0, min(k1, nz) - k0)
    assert own.min() == 1 and own.max() == 1
    assert yee == nx * ny * nz
    return tiles


def hetero_boxes(n=1024, L=12, disp_fraction=0.2):
    """The benchmark workload's boxes: TF/SF faces one cell inside the CPML
    (E faces at the box's first and last cell, H faces one below the first),
    the pole box of the dispersive slab."""
    lo, hi = L, n - L
    blo, bhi = lo + 1, hi - 1
    boxes = []
    for a in range(3):
        for plane in (blo - 1, blo, bhi):
            b = [blo, blo, blo, bhi, bhi, bhi, 4]
            b[a] = b[3 + a] = plane
            boxes.append(b)
    inner = hi - lo
    thick = int(round(disp_fraction * n ** 3 / (inner - 4) ** 2))
    z0 = max(blo + 2, (lo + hi) // 2 - thick // 2)
    boxes.append([blo + 2, blo + 2, z0, bhi - 2, bhi - 2, min(bhi - 2, z0 + thick), 2])
    return boxes


def test_plan_benchmark_geometry():
    n, L = (1024, 1024, 1024), [[12, 12]] * 3
    rows = [16] * 7 + [4]
    tiles = check(n, L, 16, 16, hetero_boxes(), rows=rows, vrows=8)
    owned = {}
    for t in tiles:
        c = int(t[0])
        cells = (min(t[1] + t[2], 1024) - t[1]) * (min(t[3] + t[4], 1024) - t[3]) * (min(t[6], 1024) - t[5])
        owned[c] = owned.get(c, 0) + cells
    cpml = sum(v for c, v in owned.items() if c < 8 and c & 1)
    # the CPML classes own the slabs and the faces next to them, not whole
    # 63-column tiles: well under the 146 M cells of a uniform 63 x 15 grid
    assert cpml < 100e6, owned
    assert owned.get(8, 0) > 600e6, owned


@pytest.mark.parametrize("R,kc,vrows", [(16, 16, 8), (8, 32, 8), (4, 7, 4)])
def test_plan_ragged(R, kc, vrows):
    rng = np.random.default_rng(20 + R)
    for trial in range(6):
        n = tuple(int(v) for v in rng.integers(9, 300, 3))
        L = [[int(rng.integers(0, 9)), int(rng.integers(0, 9))] for _ in range(3)]
        boxes = []
        for _ in range(int(rng.integers(0, 5))):
            lo = [int(rng.integers(0, n[a] + 1)) for a in range(3)]
            hi = [min(n[a], lo[a] + int(rng.integers(0, 40))) for a in range(3)]
            boxes.append(lo + hi + [int(rng.choice([2, 4]))])
        rows = [R] * 8
        rows[7] = 4 if R == 16 else R
        check(n, L, R, kc, boxes, rows=rows, vrows=vrows)
        check(n, L, R, kc, boxes, rows=rows, vec=0)


def test_plan_cavity_no_cpml():
    check((64, 64, 64), [[0, 0]] * 3, 16, 16, [[30, 30, 30, 30, 30, 30, 4]])
    check((5, 3, 2), [[0, 0]] * 3, 4, 16, [])


def test_plan_thin_x_slab_keeps_four_cell_tiles():
    """A 4-cell x CPML slab is narrower than a sliver and folds into the
    interior piece next to it; the four-cell tiles of that piece start past
    the slab (one-cell CPML tiles take its columns) instead of the whole
    piece falling back to one-cell tiles."""
    n, L = (270, 40, 36), [[4, 5], [8, 8], [8, 9]]
    boxes = [[248, 22, 20, 248, 22, 20, 4], [249, 22, 19, 249, 22, 19, 4]]
    for R, kc in ((16, 16), (8, 7), (4, 32)):
        tiles = check(n, L, R, kc, boxes, vrows=R)
        vec = tiles[tiles[:, 0] == 8]
        assert len(vec) > 0, (R, kc)
        assert vec[:, 1].min() >= 4
