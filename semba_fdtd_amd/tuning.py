"""Autotuner for the face-pass work decomposition of one grid.

The fused E and H passes are tiled as blocks of 64 x rows threads that march
zchunk z planes, each lane owning xcells consecutive x cells.  The best choice
depends on the grid's shape (ragged x edges, thin z), on binary16 storage and
on the share of dispersive cells, so ``autotune`` times each candidate on the
grid itself with device events (``fdtd_time_halves``) and keeps the fastest.
The grid's state is checkpointed before the sweep and restored after it, so
tuning leaves the fields, CPML and pole state as they were.

Results are cached per (shape, mode, storage) for the life of the process, and
optionally in a JSON file that later processes read.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Tuple

Tiling = Tuple[int, int, int]          # rows, zchunk, xcells

DEFAULT_CANDIDATES: Tuple[Tiling, ...] = tuple(
    (r, k, x) for x in (1, 4) for r in (1, 2, 4) for k in (16, 32, 64))

_cache: Dict[str, Tiling] = {}


def _key(solver) -> str:
    nx, ny, nz = solver.n3
    fp16 = int(getattr(solver, "storage", "fp32") == "fp16")
    return f"{nx}x{ny}x{nz}:{solver.mode}:{'fp16' if fp16 else 'fp32'}"


def _load(path: Optional[str]) -> None:
    if path and os.path.exists(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                _cache.setdefault(k, tuple(v))


def _save(path: Optional[str]) -> None:
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump({k: list(v) for k, v in _cache.items()}, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def autotune(solver, candidates: Iterable[Tiling] = DEFAULT_CANDIDATES, reps: int = 4,
             cache: Optional[str] = None, verbose: bool = False,
             keep_state: bool = True) -> dict:
    """Time every candidate tiling on ``solver``'s grid, set the fastest and
    return {"tiling": (rows, zchunk, xcells), "ms": E+H ms, "table": [...]}.

    Binary16 storage always runs four cells per lane, so its candidates with
    one cell per lane are skipped.  ``keep_state=False`` skips the checkpoint
    around the sweep: right for a grid at rest (all fields and pole state
    zero, which the source-free passes leave zero), where copying the state
    out and back would cost more than the sweep."""
    _load(cache)
    key = _key(solver)
    if key in _cache:
        solver.set_tiling(*_cache[key])
        return {"tiling": _cache[key], "ms": None, "table": [], "cached": True}
    fp16 = getattr(solver, "storage", "fp32") == "fp16"
    state = solver.save_state() if keep_state else None
    before = solver.get_tiling()
    table = []
    try:
        for rows, zchunk, xcells in candidates:
            if fp16 and xcells != 4:
                continue
            solver.set_tiling(rows, zchunk, xcells)
            e, h, ade = solver.time_halves(reps)
            table.append(((rows, zchunk, xcells), e, h, ade))
            if verbose:
                print(f"tiling rows={rows} zchunk={zchunk} xcells={xcells}: "
                      f"E {e:.3f} ms (poles {ade:.3f}) H {h:.3f} ms", flush=True)
    finally:
        if state is not None:
            solver.load_state(state)
    if not table:
        solver.set_tiling(*before)
        return {"tiling": before, "ms": None, "table": [], "cached": False}
    best = min(table, key=lambda r: r[1] + r[2])
    solver.set_tiling(*best[0])
    _cache[key] = best[0]
    _save(cache)
    return {"tiling": best[0], "ms": best[1] + best[2], "table": table, "cached": False}
