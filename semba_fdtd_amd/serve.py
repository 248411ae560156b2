"""Serving independent solves on one GPU: continuous batching against the
plain one-solve-at-a-time path.

A request is (arrival time, Problem, steps).  ``serve_batched`` admits
requests as they arrive while the device-memory budget allows, keeps their
fields and CPML/pole state resident, and advances all admitted grids together
(``Batch``); a grid leaves the batch as soon as its steps are done, its probe
series and frequency-domain sums are copied out, and waiting requests take
its place.  ``serve_plain`` solves the same stream one request at a time.
Both report per-request latency (arrival to results on the host) and the
aggregate Mcell-updates/s.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List

import numpy as np

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

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def device_bytes_estimate(pb) -> int:
    """Device memory of one resident grid before it is created: fields and
    media indices over the padded grid, plus CPML and pole state."""
    px = (32 + pb.nx + 2 + 31) // 32 * 32
    plane = px * (pb.ny + 3)
    base = plane * (pb.nz + 3) * (6 * 4 + 6 * 2)
    return int(base * 1.25) + (1 << 20)


def _finish(s: Solver, arrival: float, t0: float, steps: int) -> Result:
    probes = s.probes()
    dft = s.dft_sums()
    pb = s.pb
    return Result(pb.name, pb.cells, steps, arrival, time.perf_counter() - t0, probes, dft)


def serve_plain(requests, mode: str = "fast") -> List[Result]:
    """One request at a time, in arrival order, each on its own graph."""
    t0 = time.perf_counter()
    out = []
    for arrival, pb, ns in requests:
        now = time.perf_counter() - t0
        if now < arrival:
            time.sleep(arrival - now)
        s = Solver(pb, mode=mode, probe_buffer=max(ns, 1))
        s.run(ns)
        out.append(_finish(s, arrival, t0, ns))
        s.close()
    return out


def serve_batched(requests, mode: str = "fast", mem_budget: int = 200 << 30,
                  max_chunk: int = 250, rows: int = 4, zchunk: int = 32,
                  blocks: int = 0, xcells: int = 1,
                  max_cells: int | None = None) -> List[Result]:
    """Continuous batching: admit arrived requests, in arrival order, while
    the memory budget allows and (max_cells) the admitted grids hold fewer
    cells than max_cells; advance the batch in chunks that end when the next
    admitted grid finishes (or after max_chunk steps), retire finished grids.
    max_cells trades latency for throughput: once the batch holds enough
    cells to fill the device, admitting more only stretches every admitted
    request's time."""
    t0 = time.perf_counter()
    batch = Batch(rows=rows, zchunk=zchunk, blocks=blocks, xcells=xcells)
    waiting = list(requests)
    active = {}          # slot -> (solver, arrival, steps, bytes)
    used = 0
    cells = 0
    out = []
    while waiting or active:
        now = time.perf_counter() - t0
        # admission by device memory, in arrival order
        while waiting and waiting[0][0] <= now:
            arrival, pb, ns = waiting[0]
            need = device_bytes_estimate(pb)
            if active and used + need > mem_budget:
                break
            if active and max_cells is not None and cells >= max_cells:
                break
            waiting.pop(0)
            s = Solver(pb, mode=mode, probe_buffer=max(min(ns, max_chunk), 1))
            slot = batch.admit(s)
            active[slot] = (s, arrival, ns, need)
            used += need
            cells += pb.cells
        if not active:
            if waiting:
                time.sleep(max(0.0, waiting[0][0] - (time.perf_counter() - t0)))
            continue
        left = min(ns - s.n for s, _, ns, _ in active.values())
        k = max(1, min(left, max_chunk))
        batch.step(k)
        for slot in [sl for sl, (s, _, ns, _) in active.items() if s.n >= ns]:
            s, arrival, ns, need = active.pop(slot)
            batch.retire(slot)
            out.append(_finish(s, arrival, t0, ns))
            cells -= s.pb.cells
            s.close()
            used -= need
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
