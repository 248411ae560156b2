"""fdtd_get_field_noalloc, the field read that allocates no device memory,
against fdtd_get_field on the same grid: the same fp32 values bit for bit,
for every component, on two calls in a row, the second on a smaller grid (the
device array it stages through keeps the first call's contents).  FP32 and
binary16 storage; a grid large enough to take several staged runs of planes;
a component whose plane exceeds the staging array is refused before any
launch.  Called through the C ABI (ctypes)."""
import ctypes as C

import numpy as np
import pytest

import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P
from semba_fdtd_amd.solver import alloc_extent

pytestmark = pytest.mark.gpu

f32p = C.POINTER(C.c_float)


def _read_both(s, comp):
    ex, ey, ez = alloc_extent(comp, s.n3)
    a = np.full((ez, ey, ex), np.nan, np.float32)
    b = np.full((ez, ey, ex), np.nan, np.float32)
    assert s.lib.fdtd_get_field(s.h, comp, a.ctypes.data_as(f32p)) == 0
    rc = s.lib.fdtd_get_field_noalloc(s.h, comp, b.ctypes.data_as(f32p))
    assert rc == 0, s.lib.fdtd_last_error(s.h)
    return a, b


def _stepped(pb, storage, steps):
    s = sf.Solver(pb, mode="fast", storage=storage, probe_buffer=steps)
    s.run(steps)
    s.synchronize()
    return s


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_noalloc_read_equals_get_field_twice(storage):
    big = P.scatterer_problem(n=(40, 36, 44), nsteps=60)
    small = P.scatterer_problem(nsteps=60)          # 24 x 22 x 20
    assert small.cells < big.cells
    for pb in (big, small):
        print(f"noalloc read {storage} {pb.nx}x{pb.ny}x{pb.nz}", flush=True)
        s = _stepped(pb, storage, 60)
        try:
            nonzero = 0
            for comp in range(6):
                a, b = _read_both(s, comp)
                assert np.isfinite(a).all()
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (pb.name, comp)
                nonzero += int(np.count_nonzero(a))
            assert nonzero > 0
        finally:
            s.close()


def _raw_grid(lib, n, storage):
    from semba_fdtd_amd._lib import FdtdGridDesc
    d = FdtdGridDesc(n[0], n[1], n[2], 0, 1 if storage == "fp16" else 0, 0, None)
    h = C.c_void_p()
    assert lib.fdtd_create(C.byref(d), C.byref(h)) == 0, lib.fdtd_last_error(None)
    return h


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_noalloc_read_several_staged_runs(storage):
    """Random fields written with fdtd_set_field into a 1200 x 1200 x 40 grid
    (an Ex plane is 1202 x 1203 fp32 values, 5.8 MB: the 16 MiB staging array
    holds two planes, so a read takes 21 runs), read back both ways; then the
    same on a 300 x 200 x 30 grid."""
    lib = sf.load_library()
    rng = np.random.default_rng(7)
    for n in ((1200, 1200, 40), (300, 200, 30)):
        print(f"noalloc read {storage} {n[0]}x{n[1]}x{n[2]}", flush=True)
        h = _raw_grid(lib, n, storage)
        try:
            for comp in (0, 5):
                ex, ey, ez = alloc_extent(comp, n)
                src = rng.standard_normal((ez, ey, ex)).astype(np.float32)
                assert lib.fdtd_set_field(h, comp, src.ctypes.data_as(f32p)) == 0
                a = np.full_like(src, np.nan)
                b = np.full_like(src, np.nan)
                assert lib.fdtd_get_field(h, comp, a.ctypes.data_as(f32p)) == 0
                rc = lib.fdtd_get_field_noalloc(h, comp, b.ctypes.data_as(f32p))
                assert rc == 0, lib.fdtd_last_error(h)
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (n, comp)
                if storage == "fp32":
                    assert np.array_equal(a, src)
                else:
                    assert np.abs(a - src).max() <= 2e-3 * np.abs(src).max()
        finally:
            assert lib.fdtd_destroy(h) == 0


def test_noalloc_read_refuses_plane_beyond_scratch():
    """A 2100 x 2100 x 2 grid: one z plane of Ex is 2102 x 2103 fp32 values
    (17.7 MB), more than the 16 MiB staging array: refused with FDTD_ERR_ARG
    before any launch."""
    lib = sf.load_library()
    h = _raw_grid(lib, (2100, 2100, 2), "fp32")
    try:
        out = np.zeros(1, np.float32)
        rc = lib.fdtd_get_field_noalloc(h, 0, out.ctypes.data_as(f32p))
        assert rc == -1
        assert b"stages at most" in lib.fdtd_last_error(h)
    finally:
        assert lib.fdtd_destroy(h) == 0
