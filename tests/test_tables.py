"""Coefficient tables: the library's host builders against the oracle's
restatement and against the reference's own modules (oracle/_ref)."""
import numpy as np
import pytest

import oracle
import semba_fdtd_amd as sf
from semba_fdtd_amd import problem as P
from conftest import needs_ref
from refrun import run_reference

CASES = [
    ("cavity_problem", dict(n=12, nsteps=20)),
    ("scatterer_problem", dict(nsteps=20)),
    ("scatterer_problem", dict(nsteps=20, graded_z=1.03)),
    ("scatterer_problem", dict(nsteps=20, n=(30, 30, 30), L=((10, 10), (10, 10), (10, 10)),
                               alphamax=0.1, kappamax=3.0, order=2.0)),
    ("scatterer_problem", dict(nsteps=20, n=(26, 24, 22), L=((5, 0), (0, 6), (4, 4)),
                               refl=1e-3)),
]


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("builder,kw", CASES)
def test_library_tables_equal_oracle(builder, kw):
    pb = getattr(P, builder)(**kw)
    mine = sf.build_tables(pb)
    orc = oracle.oracle_tables(oracle.load_oracle(), pb)
    for name in ("g1", "g2", "gm1", "gm2", "kappa", "beta", "g3"):
        assert _same(getattr(mine, name), getattr(orc, name)), name
    for name in ("ide", "idh", "pbe", "pce", "pbm", "pcm"):
        for a in range(3):
            assert _same(getattr(mine, name)[a], getattr(orc, name)[a]), (name, a)


@needs_ref
@pytest.mark.parametrize("builder,kw", CASES)
def test_tables_equal_reference(builder, kw, tmp_path):
    pb = getattr(P, builder)(**kw)
    mine = sf.build_tables(pb)
    ref = run_reference(builder, kw, 0, str(tmp_path))["tables"]
    for got, want in zip([mine.g1, mine.g2, mine.gm1, mine.gm2], ref[:4]):
        assert _same(got, want)
    # the reference holds Idxh on [-1, n] only
    for a in range(3):
        assert _same(mine.ide[a], ref[4 + a])
        assert _same(mine.idh[a][:-1], ref[7 + a][:-1])


def test_courant_dt_matches_real4_formula():
    lib = sf.load_library()
    for d, cfl in [((0.01, 0.01, 0.01), 0.9), ((0.01, 0.02, 0.005), 0.7)]:
        got = lib.fdtd_courant_dt(*d, P.EPS0, P.MU0, cfl, 1, 0.0)
        orc = oracle.load_oracle().orc_dt(*d, P.EPS0, P.MU0, cfl, 1, 0.0)
        assert got == orc == P.courant_dt(d, cfl=cfl)
    # without -cfl, the input dt is clipped to 0.8 of the Courant limit
    big = lib.fdtd_courant_dt(0.01, 0.01, 0.01, P.EPS0, P.MU0, 1.0, 0, 1.0)
    assert big == P.courant_dt((0.01, 0.01, 0.01), dt_in=1.0)
    assert lib.fdtd_courant_dt(0.01, 0.01, 0.01, P.EPS0, P.MU0, 1.0, 0, 1e-13) == 1e-13


def test_courant_dt_closed_form():
    # dtlay = dx/(c sqrt 3) for a cube: 1.9258331e-11 s at 1 cm, 0.9 of it
    dt = P.courant_dt((0.01, 0.01, 0.01), cfl=0.9)
    c0 = 1.0 / np.sqrt(8.8541878176e-12 * 1.2566370614e-6)
    assert abs(dt - 0.9 * 0.01 / (c0 * np.sqrt(3.0))) / dt < 1e-6
    assert np.float32(dt) == np.float32(1.7332498e-11)


def test_g_tables_closed_form():
    """Vacuum G2 = dt/eps0, PEC zero, and the lossy pair is the centred
    (Crank-Nicolson) discretisation of eps dE/dt + sigma E."""
    pb = P.scatterer_problem(nsteps=4, disp=False)
    t = sf.build_tables(pb)
    assert t.g1[0] == t.g2[0] == t.gm1[0] == t.gm2[0] == 0.0
    assert t.g1[1] == 1.0 and t.gm1[1] == 1.0
    assert abs(t.g2[1] - pb.dt / P.EPS0) / t.g2[1] < 1e-6
    eps, sig = 4.0 * P.EPS0, 0.01
    a = sig * pb.dt / (2 * eps)
    assert abs(t.g1[2] - (1 - a) / (1 + a)) < 1e-7
    assert abs(t.g2[2] - pb.dt / eps / (1 + a)) / t.g2[2] < 1e-6


def test_cpml_profile_closed_form():
    """b, c of the CPML: sigma(d) = sigma_max d^m, b = exp(-sigma dt/eps0),
    c = b - 1 (kappa = 1, alpha = 0), Gedney's optimum at 10 layers."""
    n, L, m, d0 = 40, 10, 3.0, 0.01
    pb = P.scatterer_problem(n=(n, n, n), L=((L, L),) * 3, order=m, nsteps=4, disp=False)
    t = sf.build_tables(pb)
    smax = 0.8 * (m + 1) / (np.sqrt(P.MU0 / P.EPS0) * d0)
    for depth in range(1, L + 1):
        i = L - depth                                   # E node in the low slab
        s = smax * (depth / L) ** m
        b = np.exp(-s * pb.dt / P.EPS0)
        assert abs(t.pbe[0][i + 1] - b) < 2e-6
        assert abs(t.pce[0][i + 1] - (b - 1) / d0) < 2e-6 / d0 + 1e-5 * abs(b - 1) / d0
    for i in range(L, n - L + 1):                       # interior: identity
        assert t.pbe[0][i + 1] == 1.0 and t.pce[0][i + 1] == 0.0


def test_evolucion_interpolates_linearly():
    lib = sf.load_library()
    ev = np.array([0.0, 1.0, 4.0, 9.0], np.float32)
    f = lambda t: lib.fdtd_evolucion(np.float32(t), np.float32(0.5), 3,
                                     ev.ctypes.data_as(sf._lib.f32p))
    assert f(0.0) == 0.0 and f(0.5) == 1.0 and f(0.75) == 2.5 and f(1.25) == 6.5
    assert f(1.5) == 0.0                      # past the table
    assert f(-0.1) == np.float32(-0.2)        # int() truncates toward zero
    o = oracle.load_oracle()
    s = oracle.OrcNodal(deltaevol=0.5, numus=3, evol=ev.ctypes.data_as(oracle._f32p))
    for t in np.linspace(-0.3, 1.7, 41):
        assert f(t) == o.orc_evolucion(np.float32(t), s)
