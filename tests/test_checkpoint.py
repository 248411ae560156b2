"""Checkpoint / resume: a run stopped at step 50, saved, loaded into a fresh
grid with the same set-up and continued equals the uninterrupted run bit for
bit (fields, CPML psi, electric and magnetic pole state, Mur planes, plane
wave, nodal sources, DFT sums)."""
import numpy as np
import pytest

import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P

pytestmark = pytest.mark.gpu


def _problem():
    pb = P.scatterer_problem(nsteps=120, mdisp=True, mur=(1, 1, 0, 1, 1, 0))
    pw = P.open_scatterer_problem(n=32, nsteps=120, spread=8, delay=20).planewaves[0]
    pw.box = (6, 18, 5, 16, 7, 15)
    pb.planewaves = [pw]
    pb.dfts = [P.DftBox(1, (6, 16, 6, 14, 8, 10), (2, 2, 1),
                        np.array([1e9, 3e9], np.float32))]
    return pb


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_resume_equals_uninterrupted(mode):
    print(f"checkpoint {mode}", flush=True)
    pb = _problem()
    ref = sf.Solver(pb, mode=mode)
    ref.run(120)
    a = sf.Solver(pb, mode=mode)
    a.run(50)
    state = a.save_state()
    head = a.probes()
    a.close()
    b = sf.Solver(pb, mode=mode)
    b.load_state(state)
    assert b.n == 50
    b.run(70)
    got = np.concatenate([head, b.probes()])
    assert np.array_equal(got.view(np.uint32), ref.probes().view(np.uint32))
    for c in range(6):
        assert np.array_equal(b.field(c).view(np.uint32), ref.field(c).view(np.uint32)), c
    for x, y in zip(b.dft_sums(), ref.dft_sums()):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_state_rejects_other_grid():
    a = sf.Solver(P.scatterer_problem(nsteps=10), mode="exact")
    state = a.save_state()
    b = sf.Solver(P.scatterer_problem(nsteps=10), mode="fast")
    with pytest.raises(sf.FdtdError):
        b.load_state(state)
    c = sf.Solver(P.cavity_problem(n=12, nsteps=10), mode="exact")
    with pytest.raises(sf.FdtdError):
        c.load_state(state)
