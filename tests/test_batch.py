"""Batches of uneven grids (persistent face-pass launch, graph-replayed step,
continuous admission): every grid's probes, fields and frequency-domain sums
equal those of the same grid advanced alone, bit for bit, in both modes."""
import numpy as np
import pytest

import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P
from semba_fdtd_amd.batch import Batch

pytestmark = pytest.mark.gpu


def _problems():
    a = P.scatterer_problem(nsteps=150)                       # CPML, poles, nodal E and H
    b = P.open_scatterer_problem(n=30, L=5, radius=4, nsteps=120, spread=8, delay=30,
                                 grading=1.05, direction=(0.3, -0.5, 0.8),
                                 polar=(0.0, 0.848, 0.53))    # TF/SF, graded z
    b.dfts = [P.DftBox(0, (8, 20, 9, 18, 10, 12), (2, 1, 1),
                       np.array([0.5e9, 2e9, 3.5e9], np.float32))]
    c = P.cavity_problem(n=20, nsteps=100)                    # no CPML, hard Ez
    d = P.scatterer_problem(nsteps=90, n=(20, 18, 16), L=((3, 3), (2, 3), (3, 2)),
                            mur=(1, 0, 1, 0, 0, 1), mdisp=True)  # Mur, magnetic poles
    return [(a, 150), (b, 120), (c, 100), (d, 90)]


def _state(s):
    return s.probes(), [s.field(k) for k in range(6)], s.dft_sums()


@pytest.mark.parametrize("xcells", [1, 4])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_batch_equals_alone(mode, xcells):
    print(f"batch {mode} xcells {xcells}", flush=True)
    probs = _problems()
    alone = []
    for pb, ns in probs:
        s = sf.Solver(pb, mode=mode)
        s.run(ns)
        alone.append(_state(s))
        s.close()
    sol = [sf.Solver(pb, mode=mode) for pb, _ in probs]
    batch = Batch(rows=2, zchunk=8, blocks=64, xcells=xcells)
    sa = batch.admit(sol[0])
    sb = batch.admit(sol[1])
    batch.step(20)
    sc = batch.admit(sol[2])                       # joins while the others run
    sd = batch.admit(sol[3])
    batch.step(90)
    batch.retire(sd)
    batch.step(10)
    batch.retire(sb)
    batch.retire(sc)
    batch.step(30)
    batch.retire(sa)
    assert [s.n for s in sol] == [150, 120, 100, 90]
    for q, s in enumerate(sol):
        got = _state(s)
        want = alone[q]
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32)), q
        for k in range(6):
            assert np.array_equal(got[1][k].view(np.uint32), want[1][k].view(np.uint32)), (q, k)
        for g, w in zip(got[2], want[2]):
            assert np.array_equal(g.view(np.uint64), w.view(np.uint64)), q
    batch.close()


def test_retired_grid_continues_alone():
    pb = P.scatterer_problem(nsteps=80)
    ref = sf.Solver(pb, mode="exact")
    ref.run(80)
    s = sf.Solver(pb, mode="exact")
    batch = Batch()
    slot = batch.admit(s)
    batch.step(30)
    batch.retire(slot)
    s.run(50)                                      # back on its own graph
    assert np.array_equal(s.probes().view(np.uint32), ref.probes().view(np.uint32))
    batch.close()


@pytest.mark.parametrize("max_cells", [None, 60_000])
def test_serving_results_equal_plain(max_cells):
    """Continuous batching: requests arriving while others run give the same
    probe series and frequency-domain sums as the one-at-a-time path, also
    when admission is capped (60 000 cells: one or two grids at a time)."""
    from semba_fdtd_amd import serve
    reqs = P.serving_requests(count=5, rate=40.0, nmin=24, nmax=36, steps=(40, 120), nfreq=8)
    plain = {r.name: r for r in serve.serve_plain(reqs, mode="fast")}
    batched = {r.name: r for r in serve.serve_batched(reqs, mode="fast", max_chunk=25,
                                                      max_cells=max_cells)}
    assert set(plain) == set(batched)
    for name in plain:
        a, b = plain[name], batched[name]
        assert a.probes.shape == b.probes.shape and a.probes.shape[0] == a.steps
        assert np.array_equal(a.probes.view(np.uint32), b.probes.view(np.uint32)), name
        for x, y in zip(a.dft, b.dft):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), name
    s = serve.summarize(list(batched.values()))
    assert s["requests"] == 5 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_admitted_grid_refuses_stepping_and_destroy():
    """While a grid belongs to a batch its own stepping calls and
    fdtd_destroy are refused (FDTD_ERR_STATE) before any launch, and it cannot
    join a second batch; once retired it steps alone again.  Each refused call
    returns before touching the device, so nothing here can fault."""
    from semba_fdtd_amd._lib import FdtdError
    pb = P.scatterer_problem(nsteps=40)
    s = sf.Solver(pb, mode="fast")
    lib, h = s.lib, s.h
    batch = Batch()
    slot = batch.admit(s)
    ERR_STATE = -4
    assert lib.fdtd_run(h, s.n, 2) == ERR_STATE
    assert lib.fdtd_advance_e(h, 0) == ERR_STATE
    assert lib.fdtd_advance_h(h, 0) == ERR_STATE
    assert lib.fdtd_advance_e_yee(h) == ERR_STATE
    assert lib.fdtd_advance_h_cpml(h) == ERR_STATE
    assert lib.fdtd_advance_mur(h) == ERR_STATE
    assert lib.fdtd_update_observation(h, 0) == ERR_STATE
    assert lib.fdtd_sample_probes(h) == ERR_STATE
    assert lib.fdtd_destroy(h) == ERR_STATE
    assert b"batch" in lib.fdtd_last_error(h)
    other = Batch()
    with pytest.raises(FdtdError):
        other.admit(s)
    with pytest.raises(FdtdError):
        batch.admit(s)
    batch.step(10)
    batch.retire(slot)
    s.run(10)                                      # alone again
    assert s.n == 20 and s.probes().shape[0] == 20
    other.close()
    batch.close()
    s.close()


def test_run_checks_step_and_probe_room():
    """fdtd_run refuses a negative first step whatever the grid holds, and a
    run that would overflow the probe buffer between two reads (the samples
    would be lost), before any launch."""
    pb = P.cavity_problem(n=12, nsteps=40)        # nodal source only: no plane wave
    s = sf.Solver(pb, mode="fast", probe_buffer=10)
    lib, h = s.lib, s.h
    ERR_ARG = -1
    assert lib.fdtd_run(h, -1, 2) == ERR_ARG
    assert lib.fdtd_run(h, 0, 11) == ERR_ARG
    assert lib.fdtd_run(h, 0, 6) == 0
    assert lib.fdtd_run(h, 6, 6) == ERR_ARG      # 6 buffered + 6 > 10
    assert lib.fdtd_run(h, 6, 4) == 0
    assert lib.fdtd_sample_probes(h) == ERR_ARG  # full
    s.n = 10
    s._pending = 10
    got = s.probes()                               # reads and clears the buffer
    assert got.shape == (10, s.nprobe)
    assert lib.fdtd_run(h, 10, 10) == 0
    s.close()


@pytest.mark.parametrize("rows", [16, 8])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_batch_step_mode_equals_alone(mode, rows):
    """One launch per step for the batch (every grid's own tile plan, one
    launch per tile class over all grids, pairs of steps, an odd last step on
    the face passes), grids joining and leaving between calls of odd and even
    length: each grid equals itself advanced alone, bit for bit."""
    print(f"batch step mode {mode} rows {rows}", flush=True)
    probs = _problems()
    counts = [21 + 69 + 10 + 30, 21 + 69 + 10, 69 + 10, 69]
    alone = []
    for (pb, _), ns in zip(probs, counts):
        s = sf.Solver(pb, mode=mode)
        s.run(ns)
        alone.append(_state(s))
        s.close()
    sol = [sf.Solver(pb, mode=mode) for pb, _ in probs]
    batch = Batch(step_rows=rows, step_zchunk=8)
    sa = batch.admit(sol[0])
    sb = batch.admit(sol[1])
    batch.step(21)
    sc = batch.admit(sol[2])
    sd = batch.admit(sol[3])
    batch.step(69)
    batch.retire(sd)
    batch.step(10)
    batch.retire(sb)
    batch.retire(sc)
    batch.step(30)
    batch.retire(sa)
    assert [s.n for s in sol] == counts
    for q, s in enumerate(sol):
        got = _state(s)
        want = alone[q]
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32)), q
        for k in range(6):
            assert np.array_equal(got[1][k].view(np.uint32), want[1][k].view(np.uint32)), (q, k)
        for g, w in zip(got[2], want[2]):
            assert np.array_equal(g.view(np.uint64), w.view(np.uint64)), q
    batch.close()
    for s in sol:
        s.close()


def test_batch_step_mode_refuses_binary16():
    """One launch per step needs fp32 storage: a batch of binary16 grids in
    step mode refuses to step (before any launch) and steps on the face
    passes once step mode is off."""
    from semba_fdtd_amd._lib import FdtdError
    pb = P.scatterer_problem(nsteps=20, disp=False)
    s = sf.Solver(pb, mode="fast", storage="fp16")
    batch = Batch(step_rows=16)
    batch.admit(s)
    with pytest.raises(FdtdError):
        batch.step(4)
    batch.set_step_mode(0)
    batch.step(4)
    batch.close()
    assert s.n == 4
    s.close()
