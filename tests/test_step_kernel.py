"""One launch per step (k_step): E then H of each plane in one z march, the
state of step n read from one copy and step n+1 written to the other.  Per
cell the arithmetic is that of the two passes, so EXACT mode equals the
oracle bit for bit and FAST mode equals the two-pass FAST run bit for bit,
for every block shape, with CPML, electric and magnetic poles, Mur faces,
plane waves, nodal sources and frequency-domain sums, for even and odd step
counts."""
import numpy as np
import pytest

import oracle
import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P

pytestmark = pytest.mark.gpu

CASES = {
    "cavity_hard": lambda: P.cavity_problem(n=24, nsteps=120),
    "scatterer": lambda: P.scatterer_problem(nsteps=120),
    "graded_one_sided": lambda: P.scatterer_problem(nsteps=100, n=(70, 24, 22),
                                                    L=((5, 0), (0, 6), (4, 4)), graded_z=1.03),
    "wide_x": lambda: P.scatterer_problem(nsteps=80, n=(131, 9, 11),
                                          L=((6, 7), (2, 2), (3, 2))),
    # either side of the switch between the narrow and the wide row layout
    "rows_narrow_255": lambda: P.scatterer_problem(nsteps=61, n=(255, 14, 13),
                                                   L=((6, 5), (3, 3), (3, 3))),
    "rows_wide_256": lambda: P.scatterer_problem(nsteps=61, n=(256, 14, 13),
                                                 L=((6, 5), (3, 3), (3, 3))),
    "planewave_oblique": lambda: P.open_scatterer_problem(
        n=30, L=5, radius=4, nsteps=101, spread=8, delay=30, grading=1.05,
        direction=(0.3, -0.5, 0.8), polar=(0.0, 0.848, 0.53)),
    "magnetic_mur": lambda: P.scatterer_problem(nsteps=101, L=((3, 4), (2, 3), (4, 2)),
                                                mur=(1, 1, 0, 1, 1, 0), mdisp=True),
    "mur_walls": lambda: P.scatterer_problem(nsteps=90, L=((0, 0), (0, 0), (0, 0)),
                                             mur=(1, 1, 1, 1, 1, 1)),
}
SHAPES = [(4, 32), (8, 7), (4, 3), (8, 64), (16, 5)]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("name", list(CASES))
def test_step_exact_equals_oracle(name):
    pb = CASES[name]()
    ns = pb.nsteps_max
    ref = oracle.Oracle(pb)
    want = ref.run(ns)
    for rows, kc in SHAPES:
        print(f"step case {name} rows {rows} zchunk {kc}", flush=True)
        s = sf.Solver(pb, mode="exact")
        s.set_step_mode(rows, kc)
        assert s.get_step_mode() == (rows, kc)
        s.run(ns)
        got = s.probes()
        assert np.array_equal(_bits(got), _bits(want)), (rows, kc)
        for c in range(6):
            assert np.array_equal(_bits(s.field(c)), _bits(ref.fields[c])), (rows, kc, c)
        s.close()


@pytest.mark.parametrize("name", ["scatterer", "planewave_oblique", "magnetic_mur"])
def test_step_fast_equals_two_pass(name):
    pb = CASES[name]()
    ns = pb.nsteps_max
    a = sf.Solver(pb, mode="fast")
    a.run(ns)
    for rows, kc in SHAPES[:2]:
        print(f"step fast {name} rows {rows} zchunk {kc}", flush=True)
        b = sf.Solver(pb, mode="fast")
        b.set_step_mode(rows, kc)
        b.run(ns)
        assert np.array_equal(_bits(b.probes()), _bits(a.probes())), (rows, kc)
        for c in range(6):
            assert np.array_equal(_bits(b.field(c)), _bits(a.field(c))), (rows, kc, c)
        b.close()


def test_step_with_dft_and_chunks():
    """Frequency-domain sums and probes across run() calls of odd and even
    lengths equal the two-pass run."""
    pb = CASES["planewave_oblique"]()
    pb.dfts = [P.DftBox(0, (8, 20, 9, 18, 10, 12), (2, 1, 1),
                        np.array([0.5e9, 2e9, 3.5e9], np.float32)),
               P.DftBox(1, (6, 16, 6, 14, 8, 10), (2, 2, 1), np.array([1e9], np.float32))]
    ns = 100
    a = sf.Solver(pb, mode="exact")
    a.run(ns)
    b = sf.Solver(pb, mode="exact")
    b.set_step_mode(8, 16)
    for k in (7, 30, 1, 2, 60):
        b.run(k)
    assert b.n == ns
    assert np.array_equal(_bits(b.probes()), _bits(a.probes()))
    for x, y in zip(b.dft_sums(), a.dft_sums()):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_step_after_set_field():
    """Fields set from outside refresh the second copy: a grid stepped in
    step mode, then given new fields, follows the two-pass run from them
    (a lossless cavity: no psi, pole or Mur state carries over)."""
    pb = CASES["cavity_hard"]()
    ns = 60
    rng = np.random.default_rng(3)
    init = [rng.standard_normal(pb.alloc_shape).astype(np.float32) * 0.01 for _ in range(6)]
    a = sf.Solver(pb, mode="exact")
    for c in range(6):
        a.set_field(c, init[c])
    a.run(ns)
    b = sf.Solver(pb, mode="exact")
    b.set_step_mode(4, 16)
    b.run(10)                       # the second copy now holds other values
    head = b.probes().shape[0]
    for c in range(6):
        b.set_field(c, init[c])
    b.n = 0                         # the same steps again from the set fields
    b.run(ns)
    assert np.array_equal(_bits(b.probes()[head:]), _bits(a.probes()))
    for c in range(6):
        assert np.array_equal(_bits(b.field(c)), _bits(a.field(c))), c


def test_step_checkpoint_resume():
    """A run checkpointed after 41 step-mode steps and resumed in step mode
    equals the uninterrupted two-pass run (psi, poles, Mur planes)."""
    pb = CASES["magnetic_mur"]()
    ns = 90
    a = sf.Solver(pb, mode="exact")
    a.run(ns)
    b = sf.Solver(pb, mode="exact")
    b.set_step_mode(4, 16)
    b.run(41)
    state = b.save_state()
    head = b.probes()
    c3 = sf.Solver(pb, mode="exact")
    c3.set_step_mode(8, 32)
    c3.load_state(state)
    c3.run(ns - 41)
    for c in range(6):
        assert np.array_equal(_bits(c3.field(c)), _bits(a.field(c))), c
    assert np.array_equal(_bits(np.concatenate([head, c3.probes()])), _bits(a.probes()))


@pytest.mark.parametrize("name", ["graded_one_sided", "planewave_oblique", "mur_walls"])
def test_step_classes_cover_the_grid(name):
    """The tile classes own every Yee cell once, timing them leaves the grid's
    state alone, and the run after timing equals the run without it."""
    pb = CASES[name]()
    a = sf.Solver(pb, mode="fast")
    a.set_step_mode(8, 7)
    a.run(20)
    cl = a.step_classes(reps=3)
    assert sum(c[1] for c in cl) == pb.nx * pb.ny * pb.nz
    assert len({c[0] for c in cl}) == len(cl) and all(c[2] > 0 for c in cl)
    a.run(21)
    b = sf.Solver(pb, mode="fast")
    b.set_step_mode(8, 7)
    b.run(41)
    for c in range(6):
        assert np.array_equal(_bits(a.field(c)), _bits(b.field(c))), c
    assert np.array_equal(_bits(a.probes()), _bits(b.probes()))


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_step_after_set_field_with_poles(mode):
    """Fields set from outside after some steps leave the stored previous
    fields of the pole updates behind the fields: the step mode reads them
    for the next pair and then equals the two-pass run again (electric and
    magnetic poles, CPML, Mur)."""
    pb = CASES["magnetic_mur"]()
    rng = np.random.default_rng(5)
    init = [rng.standard_normal(pb.alloc_shape).astype(np.float32) * 0.01 for _ in range(6)]
    runs = []
    for step_mode in (False, True):
        s = sf.Solver(pb, mode=mode)
        if step_mode:
            s.set_step_mode(8, 7)
        s.run(12)
        for c in range(6):
            s.set_field(c, init[c])
        s.run(30)
        runs.append(s)
    a, b = runs
    for c in range(6):
        assert np.array_equal(_bits(b.field(c)), _bits(a.field(c))), c
    assert np.array_equal(_bits(b.probes()), _bits(a.probes()))


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_step_many_media(mode):
    """More media than an 8-bit index holds (300): the step kernels read the
    int16 media arrays and equal the two-pass run bit for bit (with at most
    256 media they read 8-bit copies, which every other case here covers)."""
    n, nm = 24, 300
    base = P.cavity_problem(n=n, nsteps=80)
    rng = np.random.default_rng(9)
    mi = base.mi.copy()
    inner = (slice(4, n - 3),) * 3
    for c in range(6):
        mi[c][inner] = rng.integers(2, nm + 1, size=mi[c][inner].shape).astype(mi.dtype)
    pb = P.Problem(n, n, n, epr=[1.0, 1.0] + list(rng.uniform(1.0, 6.0, nm - 1)),
                   sigma=[0.0, 0.0] + list(rng.uniform(0.0, 0.2, nm - 1)),
                   mur=[1.0] * (nm + 1), sigmam=[0.0] * (nm + 1),
                   kind=[P.KIND_PEC] + [0] * nm, mi=mi, D=base.D, dt=base.dt,
                   nodal=base.nodal, probes=base.probes, nsteps_max=80, name="many_media")
    assert pb.nmedia + 1 > 256
    runs = []
    for step_mode in (False, True):
        s = sf.Solver(pb, mode=mode)
        if step_mode:
            s.set_step_mode(8, 7)
        s.run(60)
        runs.append(s)
    a, b = runs
    for c in range(6):
        assert np.array_equal(_bits(b.field(c)), _bits(a.field(c))), c
    assert np.array_equal(_bits(b.probes()), _bits(a.probes()))


@pytest.mark.parametrize("name", ["graded_one_sided", "wide_x", "planewave_oblique",
                                  "magnetic_mur"])
def test_step_class_rows_equal_oracle(name):
    """Tile classes run in smaller blocks over row bands of the step's tiles
    (4 and 6 rows under 16-row tiles, every class or only the interior one),
    and the half-width classes (9-16) on 4 and 8 waves, over the defaults
    (some classes on fewer rows than the step's): EXACT still equals the
    oracle bit for bit, FAST the two-pass FAST run."""
    pb = CASES[name]()
    ns = pb.nsteps_max
    ref = oracle.Oracle(pb)
    want = ref.run(ns)
    a = sf.Solver(pb, mode="fast")
    a.run(ns)
    for mode in ("exact", "fast"):
        for rows, classes in ((4, 0xFF), (6, 0xFF), (6, 0x01), (4, 0x1FE00), (8, 0x1FE00)):
            print(f"step class rows {name} {mode} 16 x 5, rows {rows} classes {classes:#x}",
                  flush=True)
            s = sf.Solver(pb, mode=mode)
            s.set_step_mode(16, 5)
            default = [16, 4] + [16] * 5 + [4, 16] + [16] * 5 + [8, 16, 4]
            assert s.get_step_class_rows() == default
            s.set_step_class_rows(rows, classes)
            assert s.get_step_class_rows() == [rows if classes >> c & 1 else default[c]
                                               for c in range(17)]
            s.run(ns)
            base = ref.fields if mode == "exact" else [a.field(c) for c in range(6)]
            assert np.array_equal(_bits(s.probes()), _bits(want if mode == "exact" else a.probes()))
            for c in range(6):
                assert np.array_equal(_bits(s.field(c)), _bits(base[c])), (mode, rows, c)
            cl = s.step_classes()
            assert sum(c[1] for c in cl) == pb.nx * pb.ny * pb.nz
            s.close()
    a.close()


# grids wide enough in x for the four-cell kernel (k_step4): interior runs of
# more than 126 columns, ending in a ragged tile, and CPML slabs thick enough
# in y and z (8-9 layers) that the planner keeps them as pieces of their own,
# so interior rows and planes hold plain tiles; with CPML (the tiles stop at
# the slabs), without CPML (they reach the grid's last columns), with Mur
# faces, and with more media than 8 bits hold (int16 media reads)
def _wide_many_media():
    nx, ny, nz, nm = 270, 40, 36, 300
    base = P.scatterer_problem(nsteps=60, n=(nx, ny, nz), L=((4, 5), (8, 8), (8, 9)))
    rng = np.random.default_rng(11)
    mi = base.materialize() if base.mi is None else base.mi.copy()
    inner = (slice(3, nz - 2), slice(3, ny - 2), slice(6, nx - 6))
    for c in range(6):
        mi[c][inner] = rng.integers(2, nm + 1, size=mi[c][inner].shape).astype(mi.dtype)
    pb = P.Problem(nx, ny, nz, epr=[1.0, 1.0] + list(rng.uniform(1.0, 6.0, nm - 1)),
                   sigma=[0.0, 0.0] + list(rng.uniform(0.0, 0.2, nm - 1)),
                   mur=[1.0] * (nm + 1), sigmam=[0.0] * (nm + 1),
                   kind=[P.KIND_PEC] + [0] * nm, mi=mi, D=base.D, dt=base.dt,
                   cpml_layers=base.cpml_layers, cpml_order=base.cpml_order,
                   cpml_refl=base.cpml_refl, nodal=base.nodal, probes=base.probes,
                   nsteps_max=60, name="wide_many_media")
    assert pb.nmedia + 1 > 256
    return pb


VEC_CASES = {
    "vec_scatterer": lambda: P.scatterer_problem(nsteps=60, n=(300, 48, 44),
                                                 L=((5, 6), (8, 9), (9, 8))),
    "vec_mur_walls": lambda: P.scatterer_problem(nsteps=50, n=(270, 10, 9),
                                                 L=((0, 0), (0, 0), (0, 0)), mur=(1,) * 6),
    "vec_no_cpml": lambda: P.scatterer_problem(nsteps=80, n=(300, 40, 36),
                                               L=((0, 0), (0, 0), (0, 0)), disp=False),
    "vec_many_media": _wide_many_media,
}
VEC_SHAPES = [(16, 16), (8, 7), (4, 32)]


@pytest.mark.parametrize("name", list(VEC_CASES))
def test_step4_exact_equals_oracle(name):
    """The plain tiles run four x cells per lane (class 8) and EXACT still
    equals the oracle bit for bit, for every block shape; with the four-cell
    kernel off the same step equals it too."""
    pb = VEC_CASES[name]()
    ns = pb.nsteps_max
    ref = oracle.Oracle(pb)
    want = ref.run(ns)
    for (rows, kc), vec in [(s, True) for s in VEC_SHAPES] + [(VEC_SHAPES[0], False)]:
        print(f"step4 case {name} rows {rows} zchunk {kc} vec {vec}", flush=True)
        s = sf.Solver(pb, mode="exact")
        s.set_step_mode(rows, kc)
        s.set_step_vector(vec)
        classes = [c for c, _, _ in s.step_classes()]
        assert (8 in classes) == vec, classes
        s.run(ns)
        assert np.array_equal(_bits(s.probes()), _bits(want)), (rows, kc, vec)
        for c in range(6):
            assert np.array_equal(_bits(s.field(c)), _bits(ref.fields[c])), (rows, kc, vec, c)
        assert sum(c[1] for c in s.step_classes()) == pb.nx * pb.ny * pb.nz
        s.close()


def test_step4_fast_equals_two_pass():
    """FAST: the four-cell kernel gives the two-pass FAST run's bits."""
    pb = VEC_CASES["vec_scatterer"]()
    ns = pb.nsteps_max
    a = sf.Solver(pb, mode="fast")
    a.run(ns)
    for rows in (16, 8):
        b = sf.Solver(pb, mode="fast")
        b.set_step_mode(rows, 16)
        b.run(ns)
        assert np.array_equal(_bits(b.probes()), _bits(a.probes())), rows
        for c in range(6):
            assert np.array_equal(_bits(b.field(c)), _bits(a.field(c))), (rows, c)
        b.close()
