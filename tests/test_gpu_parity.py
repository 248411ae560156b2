"""Parity of the gfx950 path (through the C ABI) with the oracle.

EXACT mode must be bit-identical to the oracle (which is bit-identical to the
reference's own modules, tests/test_oracle_reference.py).  FAST mode (FMA
contraction, folded fp32 poles) is checked against the FP64 oracle run: its
probe relative L2 error must be <= 1e-5 (the FP32 target) or, where the
reference's own FP32 arithmetic is already further than that from FP64 (the
lossless resonant cavities), no more than 1.5x the reference's error.
"""
import numpy as np
import pytest

import oracle
import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5    # probe relative L2 vs FP64, fp32 fields
REF_FACTOR = 1.5   # allowed ratio to the reference FP32 error vs FP64


def fast_error_ok(fast, ref32, ref64):
    e_fast = rel_l2(fast.astype(np.float64), ref64)
    e_ref = rel_l2(ref32.astype(np.float64), ref64)
    print(f"rel-L2 vs FP64: fast {e_fast:.3e}, reference fp32 {e_ref:.3e}", flush=True)
    return e_fast <= max(FP32_TOL, REF_FACTOR * e_ref), (e_fast, e_ref)

CASES = {
    "cavity_hard": ("cavity_problem", dict(n=24, nsteps=300)),
    "cavity_soft": ("cavity_problem", dict(n=20, nsteps=250, hard=False)),
    "scatterer": ("scatterer_problem", dict(nsteps=300)),
    "graded": ("scatterer_problem", dict(nsteps=250, graded_z=1.03)),
    "gedney10_alpha_kappa": ("scatterer_problem", dict(nsteps=200, n=(34, 33, 32),
                                                        L=((10, 10), (10, 10), (10, 10)),
                                                        alphamax=0.1, kappamax=3.0, order=2.0)),
    "one_sided_cpml": ("scatterer_problem", dict(nsteps=200, n=(70, 24, 22),
                                                 L=((5, 0), (0, 6), (4, 4)), refl=1e-3)),
    "wide_x": ("scatterer_problem", dict(nsteps=120, n=(131, 9, 11), L=((6, 7), (2, 2), (3, 2)))),
    # the two row layouts: under 256 cells wide rows start 4 elements in with
    # a 16-byte pitch, from 256 cells 32 in with a 128-byte pitch
    "rows_narrow_255": ("scatterer_problem", dict(nsteps=120, n=(255, 14, 13),
                                                  L=((6, 5), (3, 3), (3, 3)))),
    "rows_wide_256": ("scatterer_problem", dict(nsteps=120, n=(256, 14, 13),
                                                L=((6, 5), (3, 3), (3, 3)))),
    "planewave_z": ("open_scatterer_problem", dict(n=32, L=6, radius=5, nsteps=200, spread=10,
                                                   delay=40)),
    "planewave_oblique_graded": ("open_scatterer_problem",
                                 dict(n=30, L=5, radius=4, nsteps=150, spread=8, delay=30,
                                      grading=1.05, direction=(0.3, -0.5, 0.8),
                                      polar=(0.0, 0.848, 0.53))),
    "magnetic_dispersive": ("scatterer_problem", dict(nsteps=200, mdisp=True)),
    "mur_walls": ("scatterer_problem", dict(nsteps=200, L=((0, 0), (0, 0), (0, 0)),
                                            mur=(1, 1, 1, 1, 1, 1))),
    "mur_after_cpml_mixed": ("scatterer_problem", dict(nsteps=150, L=((3, 4), (2, 3), (4, 2)),
                                                       mur=(1, 1, 0, 1, 1, 0), mdisp=True)),
}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _problem(name):
    builder, kw = CASES[name]
    return getattr(P, builder)(**kw), kw["nsteps"]


@pytest.mark.parametrize("name", list(CASES))
def test_exact_mode_bit_identical(name):
    print(f"case {name}", flush=True)
    pb, ns = _problem(name)
    ref = oracle.Oracle(pb)
    want = ref.run(ns)
    gpu = sf.Solver(pb, mode="exact")
    gpu.run(ns)
    got = gpu.probes()
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), rel_l2(got, want)
    for c in range(6):
        f = gpu.field(c)
        assert np.array_equal(f.view(np.uint32), ref.fields[c].view(np.uint32)), c
    assert np.abs(want).max() > 0


TILINGS = [(4, 32, 4), (1, 16, 4), (2, 7, 4), (2, 7, 1), (1, 64, 1), (1, 4, 1), (2, 2, 1)]


@pytest.mark.parametrize("name", list(CASES))
def test_every_tiling_bit_identical(name):
    """The fused passes with one or four x cells per lane, any rows per block
    and any z chunk (ragged ones included) give the oracle's values bit for
    bit."""
    pb, ns = _problem(name)
    ns = min(ns, 120)
    ref = oracle.Oracle(pb)
    want = ref.run(ns)
    for tiling in TILINGS:
        print(f"case {name} tiling {tiling}", flush=True)
        gpu = sf.Solver(pb, mode="exact")
        gpu.set_tiling(*tiling)
        # reported as set; a chunk past the grid's nz + 1 planes runs as one
        assert gpu.get_tiling() == tuple(tiling)
        gpu.run(ns)
        got = gpu.probes()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (tiling, rel_l2(got, want))
        for c in range(6):
            f = gpu.field(c)
            assert np.array_equal(f.view(np.uint32), ref.fields[c].view(np.uint32)), (tiling, c)
        gpu.close()


@pytest.mark.parametrize("name", ["scatterer", "one_sided_cpml", "wide_x",
                                  "planewave_oblique_graded"])
def test_fast_mode_independent_of_tiling(name):
    """FAST mode rounds with explicit fused multiply-adds, so every tiling
    (and the batch kernel) gives the same bits as the default one."""
    pb, ns = _problem(name)
    ns = min(ns, 120)
    runs = []
    for tiling in [None] + TILINGS:
        print(f"case {name} fast tiling {tiling}", flush=True)
        gpu = sf.Solver(pb, mode="fast")
        if tiling:
            gpu.set_tiling(*tiling)
        gpu.run(ns)
        runs.append((tiling, gpu.probes(), [gpu.field(c) for c in range(6)]))
        gpu.close()
    _, p0, f0 = runs[0]
    for tiling, p1, f1 in runs[1:]:
        assert np.array_equal(p1.view(np.uint32), p0.view(np.uint32)), tiling
        for c in range(6):
            assert np.array_equal(f1[c].view(np.uint32), f0[c].view(np.uint32)), (tiling, c)


def test_autotune_keeps_state():
    """Tuning checkpoints the grid and restores it: a run tuned half-way
    equals an untuned run bit for bit."""
    from semba_fdtd_amd import tuning
    pb, ns = _problem("planewave_oblique_graded")
    ref = sf.Solver(pb, mode="exact")
    ref.run(ns)
    s = sf.Solver(pb, mode="exact")
    s.run(40)
    out = tuning.autotune(s, candidates=[(4, 32, 1), (2, 16, 4), (1, 8, 4)], reps=2,
                          steps=[(4, 16), (8, 7), (16, 5)])
    assert out["tiling"] in [(4, 32, 1), (2, 16, 4), (1, 8, 4)]
    assert out["step"] in [(0, 32), (4, 16), (8, 7), (16, 5)]
    assert s.get_tiling() == out["tiling"] and s.get_step_mode() == out["step"]
    if out["class_rows"] is not None:
        assert s.get_step_class_rows() == out["class_rows"]
    assert s.n == 40
    s.run(ns - 40)
    assert np.array_equal(s.probes().view(np.uint32), ref.probes().view(np.uint32))
    for c in range(6):
        assert np.array_equal(s.field(c).view(np.uint32), ref.field(c).view(np.uint32)), c


@pytest.mark.parametrize("name", list(CASES))
def test_fast_mode_within_tolerance(name):
    print(f"case {name}", flush=True)
    pb, ns = _problem(name)
    want = oracle.Oracle(pb).run(ns)
    want64 = oracle.Oracle(pb, precision=64).run(ns)
    gpu = sf.Solver(pb, mode="fast")
    gpu.run(ns)
    ok, errs = fast_error_ok(gpu.probes(), want, want64)
    assert ok, errs


def test_cavity64_2000_steps_fast_and_exact():
    """The plain-method workload: 64^3 PEC cavity, hard Ez Gaussian source,
    2000 steps, 8 probes."""
    pb = P.cavity_problem(n=64, nsteps=2000)
    want = oracle.Oracle(pb).run(2000)
    want64 = oracle.Oracle(pb, precision=64).run(2000)
    fast = sf.Solver(pb, mode="fast")
    fast.run(2000)
    ok, errs = fast_error_ok(fast.probes(), want, want64)
    assert ok, errs
    exact = sf.Solver(pb, mode="exact")
    exact.run(2000)
    assert np.array_equal(exact.probes().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("planewave", [False, True])
def test_granular_entry_points_equal_fused(planewave):
    """advance_e_yee + advance_e_cpml + advance_e_dispersive + plane wave +
    nodal equals the fused advance_e (and the same for H), bit for bit."""
    pb = P.scatterer_problem(nsteps=60)
    if planewave:
        pw = P.open_scatterer_problem(n=32, nsteps=60, spread=8, delay=20).planewaves[0]
        pw.box = (6, 18, 5, 16, 7, 15)
        pb.planewaves = [pw]
    a = sf.Solver(pb, mode="exact")
    b = sf.Solver(pb, mode="exact")
    L = b.lib
    for n in range(60):
        a.advance_e(n)
        a.advance_h(n)
        for fn in (L.fdtd_advance_e_yee, L.fdtd_advance_e_cpml, L.fdtd_advance_e_dispersive):
            assert fn(b.h) == 0
        assert L.fdtd_advance_planewave_e(b.h, n) == 0
        assert L.fdtd_advance_nodal_e(b.h, n) == 0
        for fn in (L.fdtd_advance_h_yee, L.fdtd_advance_h_cpml):
            assert fn(b.h) == 0
        assert L.fdtd_advance_planewave_h(b.h, n) == 0
        assert L.fdtd_advance_nodal_h(b.h, n) == 0
    for c in range(6):
        assert np.array_equal(a.field(c).view(np.uint32), b.field(c).view(np.uint32)), c


@pytest.mark.parametrize("builder", ["scatterer_problem", "open_scatterer_problem"])
def test_graph_replay_equals_stepping(builder):
    pb = (P.scatterer_problem(nsteps=80) if builder == "scatterer_problem" else
          P.open_scatterer_problem(n=30, L=5, radius=4, nsteps=80, spread=8, delay=30))
    a = sf.Solver(pb, mode="fast")
    b = sf.Solver(pb, mode="fast")
    a.run(80, chunk=7)
    for _ in range(80):
        b.step()
    assert np.array_equal(a.probes().view(np.uint32), b.probes().view(np.uint32))
    for c in range(6):
        assert np.array_equal(a.field(c).view(np.uint32), b.field(c).view(np.uint32)), c


def test_device_scene_equals_host_scene():
    """fdtd_media_fill / pec_walls / paint_sphere / paint_box reproduce the
    host voxeliser (CreateVolumeMM rules) exactly; the stepped fields match
    the oracle run on the host-built matrices."""
    pb = P.heterogeneous_problem(n=72, L=6, nsteps=60, nprobes=12)
    host = P.heterogeneous_problem(n=72, L=6, nsteps=60, nprobes=12, host_media=True)
    gpu = sf.Solver(pb, mode="exact")
    ref = oracle.Oracle(host)
    want = ref.run(60)
    gpu.run(60)
    assert np.array_equal(gpu.probes().view(np.uint32), want.view(np.uint32))
    for c in range(6):
        assert np.array_equal(gpu.field(c).view(np.uint32), ref.fields[c].view(np.uint32)), c


def test_pec_walls_device_equals_host():
    n = (20, 17, 15)
    pb = P.Problem(*n, epr=[1, 1], sigma=[0, 0], mur=[1, 1], sigmam=[0, 0], kind=[1, 0],
                   mi=None, D=[P.uniform_D(v, 0.01) for v in n], dt=1e-12,
                   scene=[("fill", 1), ("pec", 0b101101)], prio=np.array([9, 0]))
    host = pb.materialize()
    gpu = sf.Solver(pb, mode="fast")
    # media leave the device only through the stepping; a field of ones on
    # every component is zeroed exactly where G1 = G2 = 0 (PEC) after one E/H
    # Yee update with zero curl
    for c in range(6):
        gpu.set_field(c, np.ones(pb.alloc_shape, np.float32))
    assert gpu.lib.fdtd_advance_e_yee(gpu.h) == 0
    assert gpu.lib.fdtd_advance_h_yee(gpu.h) == 0
    for c in range(6):
        f = gpu.field(c)
        sw = tuple(slice(1, P.sweep_hi(c, 2 - ax, n) + 2) for ax in range(3))
        assert np.array_equal(f[sw] == 0.0, host[c][sw] == 0), c


def test_media_written_between_runs():
    """The one-cell passes read no media indices in tiles whose cells all
    hold one medium, classified once per tiling and media layout.  Media
    written between two runs are followed: the lossy medium painted over a
    slab that was all vacuum (one whole tile of 4 rows x 8 planes among
    them) gives the oracle's fields with the same change, bit for bit."""
    print("media written between runs: scatterer 24x22x20, tiling 4x8x1", flush=True)
    pb = P.scatterer_problem(nsteps=120, disp=False)
    gpu = sf.Solver(pb, mode="exact")
    gpu.set_step_mode(0)
    gpu.set_tiling(4, 8, 1)
    ref = oracle.Oracle(pb)
    gpu.run(50)
    ref.run(50)
    # cells k 8..15, j 0..3, every i (Alloc index = cell index + 1)
    before = ref.mi[:, 9:17, 1:5, 1:pb.nx + 2]
    assert np.all(before == 1)
    ref.mi[:, 9:17, 1:5, 1:pb.nx + 2] = 2
    from semba_fdtd_amd._lib import i16p, ptr
    for c in range(6):
        m = P.to_alloc_box(ref.mi[c], c, gpu.n3)
        assert gpu.lib.fdtd_set_media_index(gpu.h, c, ptr(m, i16p)) == 0
    gpu.run(50)
    ref.run(50)
    for c in range(6):
        assert np.array_equal(gpu.field(c).view(np.uint32), ref.fields[c].view(np.uint32)), c


@pytest.mark.parametrize("gsteps", [1, 7])
def test_graph_steps_bit_identical(gsteps):
    """fdtd_run replaying one or several recorded two-pass steps per graph
    launch (the remainder one step per launch, over runs of uneven length;
    the default of 16 runs in every other test) gives the oracle's probes and
    fields bit for bit."""
    print(f"graph steps {gsteps}: scatterer", flush=True)
    pb, ns = _problem("scatterer")
    ref = oracle.Oracle(pb)
    want = ref.run(ns)
    gpu = sf.Solver(pb, mode="exact")
    gpu.set_step_mode(0)
    gpu.set_graph_steps(gsteps)
    for k in (5, 3 * gsteps + 2, ns - 5 - (3 * gsteps + 2)):
        gpu.run(k)
    got = gpu.probes()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), rel_l2(got, want)
    for c in range(6):
        assert np.array_equal(gpu.field(c).view(np.uint32), ref.fields[c].view(np.uint32)), c


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_full_size_linearity(mode):
    """Full-size heterogeneous grid (1024^3, 12-cell CPML, 8 lossy spheres,
    2-pole medium in 20% of the volume, TF/SF plane wave): every operation of
    the path is linear in the fields, so doubling the incident amplitude
    doubles every probe sample exactly (scaling by 2 is exact in binary
    floating point)."""
    print(f"full size linearity {mode}", flush=True)
    ns = 60
    pb = P.heterogeneous_problem(n=1024, nsteps=ns)
    w = pb.planewaves[0]
    z0 = w.box[4]
    c = pb.nx // 2
    pb.probes = [(P.EX, c, c, z0), (P.EX, c, c, z0 + 1), (P.EX, c + 7, c - 3, z0 + 2),
                 (P.HY, c, c, z0), (P.HY, c, c, z0 - 1), (P.EX, c, c, z0 - 1),
                 (P.EZ, c, c, z0)]
    a = sf.Solver(pb, mode=mode)
    a.run(ns)
    pa = a.probes()
    a.close()
    w.e = [(2.0, 0.0, 0.0)]
    b = sf.Solver(pb, mode=mode)
    b.run(ns)
    pbb = b.probes()
    b.close()
    assert np.isfinite(pa).all()
    assert np.abs(pa[:, :4]).max(axis=0).min() > 0
    # exact where the samples are normal numbers; subnormal leading-edge
    # values may round differently after doubling
    big = np.abs(pa) > 1e-30
    assert np.array_equal((2.0 * pa)[big].view(np.uint32), pbb[big].view(np.uint32))
    assert np.abs(2.0 * pa - pbb).max() <= 1e-35


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_full_size_step_equals_two_passes(mode):
    """The benchmark's path at its full size: the 1024^3 heterogeneous grid
    advanced with one launch per step (blocks of 64 x 16, 16-plane chunks,
    the default rows per tile class) gives the same probe samples, bit for
    bit, as the two passes; probes sit in the interior, the spheres, the
    pole medium, the CPML and next to the TF/SF faces."""
    print(f"full size step mode {mode}", flush=True)
    ns = 60
    pb = P.heterogeneous_problem(n=1024, nsteps=ns)
    w = pb.planewaves[0]
    z0 = w.box[4]
    c = pb.nx // 2
    pb.probes = list(pb.probes) + [(P.EX, c, c, z0), (P.EX, c, c, z0 + 1), (P.HY, c, c, z0),
                                   (P.HY, c, c, z0 - 1), (P.EZ, c, c, z0), (P.EX, 5, c, z0 + 1),
                                   (P.HY, c, 3, z0), (P.EY, c, c, 1020)]
    out = []
    for step in (False, True):
        s = sf.Solver(pb, mode=mode)
        if step:
            s.set_step_mode(16, 16)
        s.run(ns)
        out.append(s.probes())
        s.close()
    assert np.isfinite(out[0]).all() and np.abs(out[0]).max() > 0
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_dft_sums(mode):
    """Frequency-domain boxes (E and H, strided, 4 frequencies) through the
    graph and through step-wise UpdateObservation: EXACT mode equals the
    oracle's running sums bit for bit; FAST mode within the FP32 tolerance
    of the FP64 oracle."""
    print(f"dft {mode}", flush=True)
    ns = 150
    pb = P.open_scatterer_problem(n=28, L=4, radius=3, nsteps=ns, spread=8, delay=30)
    freqs = np.array([0.3e9, 1.1e9, 2.5e9, 4.0e9], np.float32)
    pb.dfts = [P.DftBox(0, (6, 20, 7, 19, 9, 14), (2, 3, 1), freqs),
               P.DftBox(1, (5, 21, 8, 12, 10, 11), (1, 2, 1), freqs[::-1].copy())]
    ref = oracle.Oracle(pb, precision=32 if mode == "exact" else 64)
    ref.run(ns)
    want = ref.dft_sums()
    a = sf.Solver(pb, mode=mode)
    a.run(ns)
    b = sf.Solver(pb, mode=mode)
    for _ in range(ns):
        b.step()
    for q in range(len(pb.dfts)):
        ga, gb = a.dft_sums()[q], b.dft_sums()[q]
        assert np.array_equal(ga.view(np.uint64), gb.view(np.uint64))
        if mode == "exact":
            assert np.array_equal(ga.view(np.uint64), want[q].view(np.uint64))
        else:
            err = np.linalg.norm(ga - want[q]) / np.linalg.norm(want[q])
            print(f"dft box {q} rel-L2 vs FP64 {err:.3e}", flush=True)
            assert err <= 1e-5
