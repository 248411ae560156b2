"""Writes the golden fixtures of tests/golden from the reference's own Fortran
modules (oracle/_ref, built by `make -C oracle ref` from the reference tree).

Each fixture holds, for one small problem of semba_fdtd_amd.problem, the probe
series (float32 [step][probe]) and the SHA-256 of each final field array in
the uniform (nz+3, ny+3, nx+3) layout.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from refrun import run_reference  # noqa: E402

CASES = {
    "cavity64_hertzian": ("cavity_problem", dict(n=64, nsteps=2000)),
    "cavity16_soft": ("cavity_problem", dict(n=16, nsteps=300, hard=False)),
    "scatterer": ("scatterer_problem", dict(nsteps=300)),
    "scatterer_graded_z": ("scatterer_problem", dict(nsteps=250, graded_z=1.03)),
    "gedney10_alpha_kappa": ("scatterer_problem", dict(nsteps=150, n=(30, 30, 30),
                                                        L=((10, 10), (10, 10), (10, 10)),
                                                        alphamax=0.1, kappamax=3.0, order=2.0)),
    "one_sided_cpml": ("scatterer_problem", dict(nsteps=150, n=(40, 20, 22),
                                                 L=((5, 0), (0, 6), (4, 4)), refl=1e-3)),
    "planewave_z_sphere": ("open_scatterer_problem", dict(n=32, L=6, radius=5, nsteps=250,
                                                          spread=10, delay=40)),
    "planewave_oblique_graded": ("open_scatterer_problem",
                                 dict(n=30, L=5, radius=4, nsteps=200, spread=8, delay=30,
                                      grading=1.05, direction=(0.3, -0.5, 0.8),
                                      polar=(0.0, 0.848, 0.53))),
    "magnetic_dispersive": ("scatterer_problem", dict(nsteps=200, mdisp=True)),
    "mur_walls": ("scatterer_problem", dict(nsteps=200, L=((0, 0), (0, 0), (0, 0)),
                                            mur=(1, 1, 1, 1, 1, 1))),
    "mur_after_cpml_mixed": ("scatterer_problem", dict(nsteps=150, L=((3, 4), (2, 3), (4, 2)),
                                                       mur=(1, 1, 0, 1, 1, 0), mdisp=True)),
}


def field_hashes(fields):
    return [hashlib.sha256(np.ascontiguousarray(f, np.float32).tobytes()).hexdigest()
            for f in fields]


def main():
    for name, (builder, kw) in CASES.items():
        with tempfile.TemporaryDirectory() as tmp:
            out = run_reference(builder, kw, kw["nsteps"], tmp)
        meta = {"builder": builder, "kwargs": kw, "nsteps": kw["nsteps"],
                "field_sha256": field_hashes(out["fields"]),
                "source": "oracle/_ref (reference Fortran modules)"}
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), probes=out["probes"],
                            meta=np.array(json.dumps(meta)))
        print(f"{name}: probes {out['probes'].shape}, |max| {np.abs(out['probes']).max():.3e}")


if __name__ == "__main__":
    main()
