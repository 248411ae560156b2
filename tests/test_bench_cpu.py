"""The benchmark's host-side contract, without a GPU: the multi-GPU request
is refused with a message, and the byte counts behind its roofline figures
reduce to their closed forms."""
import os
import subprocess
import sys
import types

import bench
from semba_fdtd_amd import problem as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_media_bytes_follow_the_one_byte_copies():
    # the library keeps one-byte copies of the media arrays when every index
    # (0..nmedia) fits a byte
    assert bench.media_bytes(types.SimpleNamespace(nmedia=255)) == 1
    assert bench.media_bytes(types.SimpleNamespace(nmedia=256)) == 2
    assert bench.onelaunch_cell_bytes(types.SimpleNamespace(nmedia=3)) == 54
    assert bench.onelaunch_cell_bytes(types.SimpleNamespace(nmedia=300)) == 60


def test_step_bytes_without_cpml_or_poles():
    """A PEC cavity has no psi and no dispersive edges: a step moves the six
    fields and the six media arrays only (two passes: each half reads three
    fields of its own, three of the other, three int16 indices, writes three)."""
    pb = P.cavity_problem(n=12, nsteps=4)
    assert bench.psi_counts(pb) == (0, 0)
    assert bench.step_algorithmic_bytes(pb, 0, "fast", 1) == 84 * pb.cells
    assert bench.onelaunch_step_bytes(pb, 0, "fast", 1) == 54 * pb.cells
    ab = bench.algorithmic_bytes(pb, 0, "fast", 1)
    assert ab["e_face"] == ab["h"] == 42 * pb.cells and ab["ade"] == 0


def test_psi_counts_of_one_slab():
    """One CPML slab of L layers on the low x face only: the psi values of
    both halves are those of the two components tangential to the slab's
    normal, over the slab's extent (closed form from the Yee staggering)."""
    n, L = (10, 7, 6), 3
    pb = types.SimpleNamespace(nx=n[0], ny=n[1], nz=n[2], cpml_layers=((L, 0), (0, 0), (0, 0)))
    nx, ny, nz = n
    e, h = bench.psi_counts(pb)
    # E: psi_Eyx over Ey edges (ny x (nz + 1)) and psi_Ezx over Ez edges
    # ((ny + 1) x nz), each L deep
    assert e == L * ny * (nz + 1) + L * (ny + 1) * nz
    # H: psi_Hyx over Hy faces ((ny + 1) x nz) and psi_Hzx over Hz faces
    # (ny x (nz + 1))
    assert h == L * (ny + 1) * nz + L * ny * (nz + 1)
