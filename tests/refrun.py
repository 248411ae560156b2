"""Runs one problem through the reference's own Fortran modules (oracle/_ref)
in a child process (their state is module-global) and returns the tables,
probe series and final fields.  Test helper."""
import os
import pickle
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_CHILD = r"""
import pickle, sys
sys.path.insert(0, {root!r})
import numpy as np
import semba_fdtd_amd.problem as P
import oracle
spec = pickle.load(open({spec!r}, 'rb'))
pb = getattr(P, spec['builder'])(**spec['kwargs'])
r = oracle.Reference(pb)
tab = r.tables()
probes = r.run(spec['nsteps'])
fields = np.stack([r.field(c) for c in range(6)])
pickle.dump({{'tables': tab, 'probes': probes, 'fields': fields}}, open({out!r}, 'wb'))
"""


def run_reference(builder: str, kwargs: dict, nsteps: int, tmpdir: str):
    spec = os.path.join(tmpdir, "spec.pkl")
    out = os.path.join(tmpdir, "out.pkl")
    with open(spec, "wb") as f:
        pickle.dump({"builder": builder, "kwargs": kwargs, "nsteps": nsteps}, f)
    code = _CHILD.format(root=ROOT, spec=spec, out=out)
    subprocess.run([sys.executable, "-c", code], check=True, cwd=tmpdir, timeout=600)
    with open(out, "rb") as f:
        return pickle.load(f)
