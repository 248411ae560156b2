"""Serving independent solves on one GPU: continuous batching against the
plain one-solve-at-a-time path.

A request is (arrival time, Problem, steps).  ``serve_batched`` admits
requests as they arrive while the device's free memory allows (hipMemGetInfo
through ``fdtd_device_memory``, less a reserve), keeps their fields and
CPML/pole state resident, and advances all admitted grids together
(``Batch``); a grid leaves the batch as soon as its steps are done, its probe
series and frequency-domain sums are copied out, and waiting requests take
its place.  ``serve_plain`` solves the same stream one request at a time.
Both report per-request latency (arrival to results on the host) and the
aggregate Mcell-updates/s.

Per-request failure handling, at every chunk boundary: a request whose probe
samples hold a non-finite value leaves the batch with status "diverged" (the
reference's own per-step NaN check is commented out, errorreport.F90:
1433-1465), and a request the caller cancels (``cancel(name)`` returns True;
the reference polls stop files, errorreport.F90:803-870) leaves with status
"cancelled".  The other requests run on unchanged.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

from . import _lib
from .batch import Batch
from .solver import Solver


@dataclass
class Result:
    name: str
    cells: int
    steps: int
    arrival: float
    done: float
    probes: np.ndarray
    dft: list
    status: str = "done"       # "done", "diverged" or "cancelled"

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def device_bytes_estimate(pb) -> int:
    """Device memory of one resident grid before it is created: fields and
    media indices over the padded grid, plus CPML and pole state."""
    px = (32 + pb.nx + 8 + 31) // 32 * 32 if pb.nx >= 256 else (4 + pb.nx + 8 + 3) // 4 * 4
    plane = px * (pb.ny + 3)
    base = plane * (pb.nz + 3) * (6 * 4 + 6 * 2)
    return int(base * 1.25) + (1 << 20)


def device_memory(device: int = 0):
    """(free, total) bytes of the device's memory (hipMemGetInfo)."""
    import ctypes as C
    lib = _lib.load()
    free, total = C.c_int64(), C.c_int64()
    rc = lib.fdtd_device_memory(device, C.byref(free), C.byref(total))
    if rc:
        raise _lib.FdtdError(f"fdtd_device_memory failed ({rc})")
    return free.value, total.value


def _finish(s: Solver, arrival: float, t0: float, steps: int, status: str = "done") -> Result:
    probes = s.probes()
    dft = s.dft_sums()
    pb = s.pb
    return Result(pb.name, pb.cells, steps, arrival, time.perf_counter() - t0, probes, dft, status)


def serve_plain(requests, mode: str = "fast", storage: str = "fp32") -> List[Result]:
    """One request at a time, in arrival order, each on its own graph."""
    t0 = time.perf_counter()
    out = []
    for arrival, pb, ns in requests:
        now = time.perf_counter() - t0
        if now < arrival:
            time.sleep(arrival - now)
        s = Solver(pb, mode=mode, probe_buffer=max(ns, 1), storage=storage)
        s.run(ns)
        out.append(_finish(s, arrival, t0, ns))
        s.close()
    return out


def serve_batched(requests, mode: str = "fast", storage: str = "fp32",
                  reserve: int = 8 << 30, max_chunk: int = 250, rows: int = 4,
                  zchunk: int = 16, blocks: int = 0, xcells: int = 1, step_rows: int = 0,
                  max_cells: Optional[int] = None, device: int = 0,
                  cancel: Optional[Callable[[str], bool]] = None,
                  check_finite: bool = True) -> List[Result]:
    """Continuous batching: admit arrived requests, in arrival order, while
    the device's free memory stays above ``reserve`` after the request's
    estimated footprint and (max_cells) the admitted grids hold fewer cells
    than max_cells; advance the batch in chunks that end when the next
    admitted grid finishes (or after max_chunk steps), retire finished grids.
    max_cells trades latency for throughput: once the batch holds enough
    cells to fill the device, admitting more only stretches every admitted
    request's time.  step_rows > 0 runs the batch one launch per step
    (``Batch.set_step_mode``).  At each chunk boundary, requests with non-finite probe
    samples (check_finite) or cancelled by the caller leave the batch early
    with status "diverged" / "cancelled"."""
    t0 = time.perf_counter()
    batch = Batch(rows=rows, zchunk=zchunk, blocks=blocks, xcells=xcells, step_rows=step_rows)
    waiting = list(requests)
    active = {}          # slot -> (solver, arrival, steps)
    cells = 0
    out = []

    def leave(slot, status):
        nonlocal cells
        s, arrival, ns = active.pop(slot)
        batch.retire(slot)
        out.append(_finish(s, arrival, t0, ns if status == "done" else s.n, status))
        cells -= s.pb.cells
        s.close()

    while waiting or active:
        now = time.perf_counter() - t0
        # admission by the device's free memory, in arrival order
        while waiting and waiting[0][0] <= now:
            arrival, pb, ns = waiting[0]
            free, _ = device_memory(device)
            # one launch per step holds a second copy of the state
            need = device_bytes_estimate(pb) * (2 if step_rows else 1)
            if active and free - need < reserve:
                break
            if active and max_cells is not None and cells >= max_cells:
                break
            waiting.pop(0)
            s = Solver(pb, mode=mode, probe_buffer=max(min(ns, max_chunk), 1), storage=storage,
                       device=device)
            slot = batch.admit(s)
            active[slot] = (s, arrival, ns)
            cells += pb.cells
        if not active:
            if waiting:
                time.sleep(max(0.0, waiting[0][0] - (time.perf_counter() - t0)))
            continue
        left = min(ns - s.n for s, _, ns in active.values())
        k = max(1, min(left, max_chunk))
        batch.step(k)
        for slot in [sl for sl, (s, _, ns) in active.items() if s.n >= ns]:
            leave(slot, "done")
        for slot in list(active):
            s = active[slot][0]
            if cancel is not None and cancel(s.pb.name):
                leave(slot, "cancelled")
            elif check_finite and s.nprobe:
                batch.synchronize()
                s._drain()
                if s._samples and not np.isfinite(s._samples[-1]).all():
                    leave(slot, "diverged")
    batch.close()
    return out


def summarize(results: List[Result], wall: float | None = None) -> dict:
    lat = np.array([r.latency for r in results])
    updates = sum(r.cells * r.steps for r in results)
    first = min(r.arrival for r in results)
    span = wall if wall is not None else max(r.done for r in results) - first
    return {"requests": len(results),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "mean_latency_s": round(float(lat.mean()), 4),
            "makespan_s": round(float(span), 3),
            "mcell_updates_per_s": round(updates / span / 1e6, 1)}
