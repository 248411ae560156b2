"""Binary16 field storage (per-array scale, fp32 arithmetic, CPML psi and
pole state in fp32) against the FP64 oracle, at the 1e-3 probe relative L2
goal.

Every stored field value is rounded to 11 significant bits once per half
step (unit roundoff 2^-11 = 4.9e-4), and the rounding errors travel with the
waves; their sum grows roughly as the square root of the step count
(rounding to nearest; stochastic rounding adds variance and was measured
worse).  Measured on the MI355X, no case meets the goal: 1.3e-3 after 100
steps of a plane wave, 2.3e-3 after the batched configuration's 1500 steps,
3.4e-3 for a serving request with a pole medium, while fp32 storage of the
same runs stays at 3e-7 to 9e-7.  Each case is marked xfail (strict: a pass
would be reported) with its measured error; the bound is the goal and is
never raised.
"""
import numpy as np
import pytest

import oracle
import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P
from semba_fdtd_amd.batch import Batch

pytestmark = pytest.mark.gpu

FP16_TOL = 1e-3   # probe relative L2 vs FP64


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _planewave_z(ns):
    return P.open_scatterer_problem(n=32, L=6, radius=5, nsteps=ns, spread=10, delay=40)


def _planewave_oblique(ns):
    return P.open_scatterer_problem(n=30, L=5, radius=4, nsteps=ns, spread=8, delay=30,
                                    grading=1.05, direction=(0.3, -0.5, 0.8),
                                    polar=(0.0, 0.848, 0.53))


def _instance(ns):
    pb = P.batch_instances()[129]          # 69 x 90 x 38, a batched-configuration instance
    pb.nsteps_max = ns
    return pb


def _request(ns):
    return P.serving_requests()[7][1]      # 62^3, 64-frequency DFT, 527 steps


CASES = {
    "planewave_z": _planewave_z,
    "planewave_oblique": _planewave_oblique,
    "instance129": _instance,
    "request7": _request,
}

# (case, steps) -> measured rel-L2 where binary16 storage misses the goal
# (MI355X, probe rel-L2 against the FP64 oracle)
OVER = {
    ("planewave_z", 100): 1.280e-3,
    ("planewave_z", 200): 1.972e-3,
    ("planewave_oblique", 150): 1.416e-3,
    ("instance129", 300): 1.336e-3,
    ("instance129", 1500): 2.276e-3,
    ("request7", 527): 3.359e-3,
}

RUNS = [("planewave_z", 100), ("planewave_z", 200), ("planewave_oblique", 150),
        ("instance129", 300), ("instance129", 1500), ("request7", 527)]


def _params():
    out = []
    for name, ns in RUNS:
        marks = []
        if (name, ns) in OVER:
            marks.append(pytest.mark.xfail(strict=True, reason=(
                f"binary16 storage: rel-L2 {OVER[(name, ns)]:.2e} vs FP64 after {ns} steps "
                f"(goal {FP16_TOL:g})")))
        out.append(pytest.param(name, ns, marks=marks, id=f"{name}-{ns}"))
    return out


@pytest.mark.parametrize("name,ns", _params())
def test_fp16_storage_against_fp64(name, ns):
    pb = CASES[name](ns)
    want64 = oracle.Oracle(pb, precision=64).run(ns)
    s = sf.Solver(pb, mode="fast", storage="fp16", probe_buffer=ns)
    s.run(ns)
    got = s.probes()
    s.close()
    f32 = sf.Solver(pb, mode="fast", probe_buffer=ns)
    f32.run(ns)
    err32 = rel_l2(f32.probes(), want64)
    f32.close()
    err = rel_l2(got, want64)
    print(f"fp16 {name} {pb.nx}x{pb.ny}x{pb.nz} {ns} steps: rel-L2 vs FP64 binary16 storage "
          f"{err:.3e}, fp32 storage {err32:.3e}", flush=True)
    assert np.isfinite(got).all()
    assert err <= FP16_TOL


@pytest.mark.parametrize("xcells", [1, 4])
def test_fp16_batch_equals_alone(xcells):
    """Binary16 grids in one batch (persistent launch, flat one-cell tiles or
    four-cell tiles) give the same bits as each grid advanced alone, in the
    one-cell and in the four-cell passes."""
    insts = [P.batch_instances()[q] for q in (2, 129, 248)]
    ns = 200
    alone = []
    for pb in insts:
        runs = []
        for xc in (1, 4):
            s = sf.Solver(pb, mode="fast", storage="fp16", probe_buffer=ns)
            s.set_tiling(4, 32, xc)
            assert s.get_tiling() == (4, 32, xc)
            s.run(ns)
            runs.append(s.probes())
            s.close()
        assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))
        alone.append(runs[0])
    sol = [sf.Solver(pb, mode="fast", storage="fp16", probe_buffer=ns) for pb in insts]
    b = Batch(xcells=xcells)
    for s in sol:
        b.admit(s)
    b.step(ns)
    b.close()
    for q, s in enumerate(sol):
        got = s.probes()
        assert np.array_equal(got.view(np.uint32), alone[q].view(np.uint32)), q
        s.close()
