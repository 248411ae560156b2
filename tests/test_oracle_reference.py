"""The oracle against the reference's own Fortran modules (oracle/_ref).

The reference harness compiles calc_constants.F90, borderscpml.F90,
electricdispersive.F90, nodalsources.F90 and planewaves.F90 from their
sources and calls them in ciclo_temporal's order (timestepping.F90:
1248-1981); the Yee loops in it restate timestepping.F90:2165-2409.  The C oracle must reproduce its probe
series and final fields bit for bit (both built without FMA contraction).
Skipped where the reference tree is absent; tests/golden holds the same
series as committed fixtures for that case (test_golden.py).
"""
import numpy as np
import pytest

import oracle
from semba_fdtd_amd import problem as P
from conftest import needs_ref
from refrun import run_reference

CASES = {
    "cavity_hard": ("cavity_problem", dict(n=16, nsteps=200)),
    "cavity_soft": ("cavity_problem", dict(n=14, nsteps=200, hard=False)),
    "scatterer": ("scatterer_problem", dict(nsteps=250)),
    "graded_z": ("scatterer_problem", dict(nsteps=200, graded_z=1.03)),
    "gedney10_alpha_kappa": ("scatterer_problem", dict(nsteps=150, n=(30, 30, 30),
                                                        L=((10, 10), (10, 10), (10, 10)),
                                                        alphamax=0.1, kappamax=3.0, order=2.0)),
    "one_sided_cpml": ("scatterer_problem", dict(nsteps=150, n=(40, 20, 22),
                                                 L=((5, 0), (0, 6), (4, 4)), refl=1e-3)),
    "hard_sources": ("scatterer_problem", dict(nsteps=150, soft=False)),
    "planewave_z": ("open_scatterer_problem", dict(n=32, L=6, radius=5, nsteps=200, spread=10,
                                                   delay=40)),
    "planewave_oblique_graded": ("open_scatterer_problem",
                                 dict(n=30, L=5, radius=4, nsteps=150, spread=8, delay=30,
                                      grading=1.05, direction=(0.3, -0.5, 0.8),
                                      polar=(0.0, 0.848, 0.53))),
    "planewave_backward": ("open_scatterer_problem",
                           dict(n=28, L=4, radius=3, nsteps=150, spread=8, delay=30,
                                direction=(-0.6, 0.0, -0.8), polar=(0.8, 0.0, -0.6),
                                margin=2)),
    # magnetic pole-residue media: a medium both electric and magnetic
    # dispersive, and a magnetic-only Lorentz block
    "magnetic_dispersive": ("scatterer_problem", dict(nsteps=150, mdisp=True)),
    "magnetic_only": ("scatterer_problem", dict(nsteps=150, mdisp=True, disp=False)),
    # first-order Mur: on every face of a grid without CPML, on top of CPML
    # faces (MurAfterPML), and mixed with CPML and bare faces
    "mur_walls": ("scatterer_problem", dict(nsteps=150, L=((0, 0), (0, 0), (0, 0)),
                                            mur=(1, 1, 1, 1, 1, 1))),
    "mur_after_cpml": ("scatterer_problem", dict(nsteps=150, mur=(1, 0, 0, 1, 0, 1))),
    "mur_mixed": ("scatterer_problem", dict(nsteps=150, L=((3, 4), (2, 3), (4, 2)),
                                            mur=(1, 1, 0, 1, 1, 0), mdisp=True)),
    "mur_partial": ("scatterer_problem", dict(nsteps=150, L=((0, 0), (0, 0), (0, 0)),
                                              mur=(1, 0, 1, 0, 0, 1))),
    # CPML on one axis only: the reference still builds all three profiles
    "cpml_x_only": ("scatterer_problem", dict(nsteps=120, L=((4, 5), (0, 0), (0, 0)))),
}


@needs_ref
@pytest.mark.parametrize("name", list(CASES))
def test_oracle_bit_identical_to_reference_modules(name, tmp_path):
    builder, kw = CASES[name]
    ns = kw["nsteps"]
    ref = run_reference(builder, kw, ns, str(tmp_path))
    pb = getattr(P, builder)(**kw)
    o = oracle.Oracle(pb)
    # coefficient tables first (G1/G2/GM1/GM2 and the stretched inverse spacings)
    t = o.t
    mine = [t.g1, t.g2, t.gm1, t.gm2] + t.ide + t.idh
    for a, b in zip(ref["tables"], mine):
        assert np.array_equal(a[:len(b) - 1].view(np.uint32), b[:len(b) - 1].view(np.uint32))
    got = o.run(ns)
    assert np.array_equal(got.view(np.uint32), ref["probes"].view(np.uint32))
    for c in range(6):
        assert np.array_equal(o.fields[c].view(np.uint32), ref["fields"][c].view(np.uint32)), c
    # the series must not be trivially zero
    assert np.abs(got).max() > 0
