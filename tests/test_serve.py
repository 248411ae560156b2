"""Serving robustness: admission by the device's free memory, a request
whose fields blow up leaves the batch as "diverged", a cancelled request
leaves as "cancelled", and every other request still gets the plain path's
results bit for bit."""
import numpy as np
import pytest

from semba_fdtd_amd import problem as P
from semba_fdtd_amd import serve

pytestmark = pytest.mark.gpu


def test_device_memory_reports_hbm():
    free, total = serve.device_memory(0)
    print(f"device memory: {free / 2**30:.1f} GiB free of {total / 2**30:.1f} GiB", flush=True)
    assert 0 < free <= total
    assert total > 100 * 2**30          # an MI355X holds 288 GB


def _stream():
    reqs = P.serving_requests(count=4, rate=50.0, nmin=24, nmax=32, steps=(120, 200), nfreq=8)
    # request 2 runs at 2.5x the Courant limit: its fields grow without bound
    t, pb, ns = reqs[2]
    pb.dt = pb.dt * 2.5
    pb.name = "unstable"
    reqs[2] = (t, pb, ns)
    return reqs


def test_diverged_and_cancelled_requests_leave_the_batch():
    reqs = _stream()
    plain = {r.name: r for r in serve.serve_plain([q for q in reqs if q[1].name != "unstable"],
                                                  mode="fast")}
    batched = {r.name: r for r in serve.serve_batched(
        reqs, mode="fast", max_chunk=20, cancel=lambda name: name == "request1")}
    assert set(batched) == {pb.name for _, pb, _ in reqs}
    assert batched["unstable"].status == "diverged"
    assert not np.isfinite(batched["unstable"].probes).all()
    assert batched["request1"].status == "cancelled"
    assert batched["request1"].probes.shape[0] < reqs[1][2]
    # what a cancelled request produced before it left equals the plain run's
    n1 = batched["request1"].probes.shape[0]
    assert np.array_equal(batched["request1"].probes.view(np.uint32),
                          plain["request1"].probes[:n1].view(np.uint32))
    for name in ("request0", "request3"):
        assert batched[name].status == "done"
        assert np.array_equal(batched[name].probes.view(np.uint32),
                              plain[name].probes.view(np.uint32)), name
