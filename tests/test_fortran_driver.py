"""The Fortran side of the boundary: the iso_c_binding module and the driver
program (semba_fdtd_amd/fortran) that keeps ciclo_temporal's call structure.

CPU: the driver reads a case file, builds every coefficient table through the
library's table builders in Fortran, and the tables equal the Python set-up's
(which the table tests pin against the oracle and the reference's own
modules).  Without a GPU it stops at fdtd_create with the library's message.
GPU: the driver's probe series equal the oracle's, bit for bit in EXACT mode,
for the fused, split (one call per reference routine) and graph loops.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P
from semba_fdtd_amd.case_file import read_probe_file, write_case

DRIVER = os.path.join(os.path.dirname(sf.__file__), "fortran", "fdtd_gpu_driver")
needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER),
                                  reason="Fortran driver not built (make -C semba_fdtd_amd/fortran)")


def _run(case, out, *args, timeout=300):
    return subprocess.run([DRIVER, case, out, *args], capture_output=True, text=True,
                          timeout=timeout)


def _read_tables(path, pb):
    nm = pb.nmedia
    with open(path, "rb") as f:
        raw = f.read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(raw, dt, n, off)
        off += a.nbytes
        return a

    g = [take("<f4", nm + 1) for _ in range(4)]
    axes = []
    for n in (pb.nx, pb.ny, pb.nz):
        axes.append([take("<f4", n + 3) for _ in range(6)])
    nd = int(take("<i4", 1)[0])
    m = max(nd, 1) * 8 * 2
    kappa, beta, g3 = (take("<f8", m).reshape(max(nd, 1), 8, 2) for _ in range(3))
    nmd = int(take("<i4", 1)[0])
    m = max(nmd, 1) * 8 * 2
    mkappa, mbeta, gm3 = (take("<f8", m).reshape(max(nmd, 1), 8, 2) for _ in range(3))
    cab1 = take("<f4", 6 * (nm + 1)).reshape(6, nm + 1)
    assert off == len(raw)
    return g, axes, (nd, kappa, beta, g3), (nmd, mkappa, mbeta, gm3), cab1


@needs_driver
@pytest.mark.parametrize("case", ["scatterer", "graded_gedney", "cavity", "magnetic_mur",
                                  "x_only_cpml"])
def test_driver_tables_equal_python_setup(tmp_path, case):
    if case == "scatterer":
        pb = P.scatterer_problem(nsteps=10)
    elif case == "magnetic_mur":
        pb = P.scatterer_problem(nsteps=10, mdisp=True, mur=(1, 1, 0, 1, 1, 0), graded_z=1.05)
    elif case == "x_only_cpml":
        pb = P.scatterer_problem(nsteps=10, L=((4, 5), (0, 0), (0, 0)), mur=(0, 0, 1, 1, 0, 0))
    elif case == "graded_gedney":
        pb = P.scatterer_problem(nsteps=10, n=(34, 33, 32), L=((10, 10), (10, 10), (10, 10)),
                                 alphamax=0.1, kappamax=3.0, order=2.0, graded_z=1.03)
    else:
        pb = P.cavity_problem(n=16, nsteps=10)
    cf, out = str(tmp_path / "case.bin"), str(tmp_path / "tables.bin")
    write_case(pb, cf, 10)
    r = _run(cf, out, "fast", "tables")
    assert r.returncode == 0, r.stderr
    g, axes, (nd, kappa, beta, g3), (nmd, mkappa, mbeta, gm3), cab1 = _read_tables(out, pb)
    t = sf.build_tables(pb)
    for got, want in zip(g, (t.g1, t.g2, t.gm1, t.gm2)):
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    for a in range(3):
        want = (t.ide[a], t.idh[a], t.pbe[a], t.pce[a], t.pbm[a], t.pcm[a])
        for got, w in zip(axes[a], want):
            assert np.array_equal(got.view(np.uint32), w.view(np.uint32)), a
    assert nd == len(t.disp_media)
    if nd:
        assert np.array_equal(kappa, t.kappa[:nd])
        assert np.array_equal(beta, t.beta[:nd])
        assert np.array_equal(g3, t.g3[:nd])
    assert nmd == len(t.mdisp_media)
    if nmd:
        assert np.array_equal(mkappa, t.mkappa[:nmd])
        assert np.array_equal(mbeta, t.mbeta[:nmd])
        assert np.array_equal(gm3, t.gm3[:nmd])
    if any(pb.mur_faces):
        assert np.array_equal(cab1.view(np.uint32), t.cab1.view(np.uint32))


@needs_driver
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_driver_without_gpu_fails_loudly(tmp_path):
    pb = P.cavity_problem(n=8, nsteps=4)
    cf, out = str(tmp_path / "case.bin"), str(tmp_path / "p.bin")
    write_case(pb, cf, 4)
    r = _run(cf, out)
    assert r.returncode != 0
    assert "fdtd_create" in r.stderr


@needs_driver
def test_driver_rejects_bad_case(tmp_path):
    cf = tmp_path / "bad.bin"
    cf.write_bytes(b"XXXX\x01\x00\x00\x00")
    r = _run(str(cf), str(tmp_path / "o.bin"), "fast", "tables")
    assert r.returncode != 0
    assert "case file" in r.stderr


GPU_CASES = {
    "scatterer": lambda: (P.scatterer_problem(nsteps=200), 200),
    "planewave_oblique_graded": lambda: (P.open_scatterer_problem(
        n=30, L=5, radius=4, nsteps=150, spread=8, delay=30, grading=1.05,
        direction=(0.3, -0.5, 0.8), polar=(0.0, 0.848, 0.53)), 150),
    "cavity64_hertzian_2000": lambda: (P.cavity_problem(n=64, nsteps=2000), 2000),
    "magnetic_mur": lambda: (P.scatterer_problem(nsteps=150, L=((3, 4), (2, 3), (4, 2)),
                                                 mur=(1, 1, 0, 1, 1, 0), mdisp=True), 150),
}


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("loop", ["fused", "split", "graph"])
@pytest.mark.parametrize("name", list(GPU_CASES))
def test_driver_exact_bit_identical(tmp_path, name, loop):
    print(f"driver case {name} loop {loop}", flush=True)
    pb, ns = GPU_CASES[name]()
    cf, out = str(tmp_path / "case.bin"), str(tmp_path / "p.bin")
    write_case(pb, cf, ns)
    r = _run(cf, out, "exact", loop)
    assert r.returncode == 0, r.stderr
    got = read_probe_file(out)
    want = oracle.Oracle(pb).run(ns)
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@needs_driver
@pytest.mark.gpu
def test_driver_fast_within_tolerance(tmp_path):
    pb, ns = P.scatterer_problem(nsteps=300), 300
    cf, out = str(tmp_path / "case.bin"), str(tmp_path / "p.bin")
    write_case(pb, cf, ns)
    r = _run(cf, out, "fast", "graph")
    assert r.returncode == 0, r.stderr
    got = read_probe_file(out).astype(np.float64)
    want = oracle.Oracle(pb, precision=64).run(ns)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"driver fast rel-L2 vs FP64 oracle {err:.3e}", flush=True)
    assert err <= 1e-5
