"""Golden fixtures written from the reference's own Fortran modules
(tests/golden/make_golden.py): probe series and SHA-256 of the final fields.

CPU: the oracle reproduces every fixture bit for bit (so the oracle stays
pinned where the reference tree is not available).  GPU: EXACT mode
reproduces them bit for bit through the C ABI.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from semba_fdtd_amd import problem as P

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(os.path.splitext(os.path.basename(f))[0]
                  for f in glob.glob(os.path.join(HERE, "*.npz")))


def load(name):
    with np.load(os.path.join(HERE, name + ".npz")) as z:
        meta = json.loads(str(z["meta"]))
        probes = z["probes"].copy()
    pb = getattr(P, meta["builder"])(**meta["kwargs"])
    return pb, meta, probes


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def test_fixtures_present():
    assert len(FIXTURES) >= 6


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_golden(name):
    pb, meta, want = load(name)
    o = oracle.Oracle(pb)
    got = o.run(meta["nsteps"])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert [sha(o.fields[c]) for c in range(6)] == meta["field_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_exact_reproduces_golden(name):
    import semba_fdtd_amd as sf
    print(f"golden {name}", flush=True)
    pb, meta, want = load(name)
    s = sf.Solver(pb, mode="exact")
    s.run(meta["nsteps"])
    got = s.probes()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert [sha(s.field(c)) for c in range(6)] == meta["field_sha256"]
