"""Autotuner for the work decomposition of one grid.

Two ways to advance a step are timed on the grid itself with device events:

* the two passes (E, then H), tiled as blocks of 64 x rows threads marching
  zchunk z planes, each lane owning xcells consecutive x cells
  (``fdtd_time_halves``);
* one launch per step (``fdtd_set_step_mode``, ``fdtd_time_step``), blocks of
  64 x rows threads marching zchunk planes, which reads every field once and
  writes it once per step but keeps a second copy of the state.  Its tile
  classes (interior, CPML, poles, sources) run kernels of their own; each
  class then gets the block rows that run its tiles fastest
  (``fdtd_set_step_class_rows``, timed per class with ``fdtd_step_classes``).

The best choice depends on the grid's shape (ragged x edges, thin z), on
binary16 storage and on the share of dispersive cells, so ``autotune`` keeps
the fastest.  The grid's state is checkpointed before the sweep and restored
after it, so tuning leaves the fields, CPML and pole state as they were.

Results are cached per (shape, mode, storage) for the life of the process, and
optionally in a JSON file that later processes read.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Tuple

from ._lib import FdtdError

Tiling = Tuple[int, int, int]          # rows, zchunk, xcells
StepShape = Tuple[int, int]            # rows, zchunk (rows 0: two passes)

# short z chunks give small grids more blocks: a 64^3 cavity steps in 22 us
# with 4-plane chunks against 41 us with 16
DEFAULT_CANDIDATES: Tuple[Tiling, ...] = tuple(
    (r, k, x) for x in (1, 4) for r in (1, 2, 4) for k in (4, 8, 16, 32, 64))
STEP_CANDIDATES: Tuple[StepShape, ...] = ((4, 16), (8, 16), (8, 32), (16, 16), (16, 32))

NCLASS = 17                            # tile classes of the one-launch step

_cache: Dict[str, tuple] = {}


def _key(solver) -> str:
    nx, ny, nz = solver.n3
    return f"{nx}x{ny}x{nz}:{solver.mode}:{getattr(solver, 'storage', 'fp32')}"


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


def _apply(solver, choice: tuple) -> None:
    rows, zchunk, xcells, srows, skc = choice[:5]
    solver.set_tiling(rows, zchunk, xcells)
    solver.set_step_mode(srows, skc)
    solver.set_step_class_rows(0)
    for c, r in enumerate(choice[5:5 + NCLASS]):
        solver.set_step_class_rows(r, 1 << c)


def _class_rows(solver, reps: int, verbose: bool) -> list:
    """Per tile class, the block rows whose kernel runs the class fastest:
    classes 0-7 among the step's rows and the row bands that divide its
    tiles, the four-cell and half-width classes among every supported block
    up to the step's rows (a request a class cannot take leaves it on its
    default, which is then timed again)."""
    srows = solver.get_step_mode()[0]
    options = [r for r in (srows, 4, 6, 8) if r <= srows]
    default = solver.get_step_class_rows()
    best = {}
    for r in dict.fromkeys(options):
        solver.set_step_class_rows(r)
        for c, _, ms in solver.step_classes(reps):
            if verbose:
                print(f"class {c} rows {r}: {ms:.3f} ms", flush=True)
            if c not in best or ms < best[c][1]:
                best[c] = (r, ms)
    rows = list(default)
    for c, (r, _) in best.items():
        rows[c] = r
    return rows


def autotune(solver, candidates: Iterable[Tiling] = DEFAULT_CANDIDATES, reps: int = 4,
             cache: Optional[str] = None, verbose: bool = False, keep_state: bool = True,
             steps: Iterable[StepShape] = STEP_CANDIDATES) -> dict:
    """Time every candidate on ``solver``'s grid and set the fastest.

    Returns {"tiling": (rows, zchunk, xcells), "step": (rows, zchunk),
    "class_rows": [rows of each tile class] or None, "ms": ms per step of the
    choice, "table": [...], "step_table": [...]}; step rows 0 means the two
    passes won.  Binary16 storage has no one-launch step.  ``keep_state=False`` skips the
    checkpoint around the sweep: right for a grid at rest (all fields and pole
    state zero, which the source-free passes leave zero), where copying the
    state out and back would cost more than the sweep."""
    _load(cache)
    key = _key(solver)
    if key in _cache:
        _apply(solver, _cache[key])
        c = _cache[key]
        return {"tiling": c[:3], "step": c[3:5], "class_rows": list(c[5:5 + NCLASS]) or None,
                "ms": None, "table": [], "step_table": [], "cached": True}
    fp16 = getattr(solver, "storage", "fp32") == "fp16"
    state = solver.save_state() if keep_state else None
    before = solver.get_tiling() + solver.get_step_mode()
    table, step_table = [], []
    class_rows = None
    try:
        solver.set_step_mode(0, 32)
        for rows, zchunk, xcells in candidates:
            solver.set_tiling(rows, zchunk, xcells)
            e, h, ade = solver.time_halves(reps)
            table.append(((rows, zchunk, xcells), e, h, ade))
            if verbose:
                print(f"tiling rows={rows} zchunk={zchunk} xcells={xcells}: "
                      f"E {e:.3f} ms (poles {ade:.3f}) H {h:.3f} ms", flush=True)
        if not fp16:
            for rows, zchunk in steps:
                try:
                    solver.set_step_mode(rows, zchunk)
                except FdtdError:
                    continue
                solver.set_step_class_rows(0)
                ms = solver.time_step(reps)
                step_table.append(((rows, zchunk), ms))
                if verbose:
                    print(f"one launch per step rows={rows} zchunk={zchunk}: {ms:.3f} ms",
                          flush=True)
        if step_table:
            bs = min(step_table, key=lambda r: r[1])
            best_passes = min((r[1] + r[2] for r in table), default=float("inf"))
            if bs[1] < best_passes:
                solver.set_step_mode(*bs[0])
                class_rows = _class_rows(solver, reps, verbose)
                solver.set_step_class_rows(0)
                for c, r in enumerate(class_rows):
                    solver.set_step_class_rows(r, 1 << c)
                ms = solver.time_step(reps)
                step_table.append((tuple(bs[0]) + ("class rows",), ms))
                if ms >= bs[1]:
                    class_rows = None
    finally:
        if state is not None:
            solver.load_state(state)
    if not table:
        _apply(solver, before)
        return {"tiling": before[:3], "step": before[3:], "class_rows": None, "ms": None,
                "table": [], "step_table": step_table, "cached": False}
    best = min(table, key=lambda r: r[1] + r[2])
    ms = best[1] + best[2]
    step = (0, 32)
    if step_table:
        bs = min(step_table, key=lambda r: r[1])
        if bs[1] < ms:
            step, ms = tuple(bs[0][:2]), bs[1]
    if step[0] == 0 or len(bs[0]) == 2:
        class_rows = None
    choice = tuple(best[0]) + tuple(step) + tuple(class_rows or ())
    _apply(solver, choice)
    _cache[key] = choice
    _save(cache)
    return {"tiling": best[0], "step": step, "class_rows": class_rows, "ms": ms,
            "table": table, "step_table": step_table, "cached": False}
