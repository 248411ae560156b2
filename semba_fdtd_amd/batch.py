"""Batches of independent grids advanced together on one GPU (the
``fdtd_batch_*`` entry points of the C ABI).

A grid is set up as a ``Solver`` and then admitted; from then on
``Batch.step`` advances every admitted grid, each from its own step, with one
persistent face-pass launch per half step for the whole batch and the step
replayed from a HIP graph.  Grids can be admitted and retired between steps
(continuous batching).  A grid's results are the same as when it is advanced
alone.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import FdtdError


class Batch:
    def __init__(self, device: int = 0, stream=None, rows: int = 4, zchunk: int = 16,
                 blocks: int = 0, xcells: int = 1, step_rows: int = 0, step_zchunk: int = 16):
        self.lib = lib = _lib.load()
        b = C.c_void_p()
        rc = lib.fdtd_batch_create(device, stream, C.byref(b))
        if rc:
            raise FdtdError(f"fdtd_batch_create failed ({rc}): "
                            f"{lib.fdtd_batch_last_error(None).decode()}")
        self.b = b
        self.members = {}          # slot -> Solver
        self.set_tiling(rows, zchunk, blocks, xcells)
        self.set_step_mode(step_rows, step_zchunk)

    def _ck(self, rc, what):
        if rc:
            msg = self.lib.fdtd_batch_last_error(self.b)
            raise FdtdError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_tiling(self, rows: int, zchunk: int, blocks: int = 0, xcells: int = 1):
        """Blocks of 64 x rows threads marching zchunk z planes, xcells x
        cells per lane (1: each block takes a run of 64 * rows consecutive
        columns of a grid's x-y plane; 4: 256-cell x segments), and the
        persistent launch's block count (0: 16384)."""
        self._ck(self.lib.fdtd_batch_set_tiling(self.b, rows, zchunk, xcells, blocks),
                 "fdtd_batch_set_tiling")

    def set_step_mode(self, rows: int, zchunk: int = 16):
        """One launch per step for the whole batch (rows 4, 8 or 16 per
        block, zchunk planes per chunk; 0: the face passes): every grid's
        own tile plan, one launch per tile class over all grids, steps in
        pairs.  fp32 storage only; the results do not change."""
        self._ck(self.lib.fdtd_batch_set_step_mode(self.b, rows, zchunk),
                 "fdtd_batch_set_step_mode")

    def admit(self, solver) -> int:
        """Admit a set-up Solver at its current step; returns its slot."""
        slot = C.c_int32()
        self._ck(self.lib.fdtd_batch_admit(self.b, solver.h, solver.n, C.byref(slot)),
                 "fdtd_batch_admit")
        self.members[slot.value] = solver
        return slot.value

    def retire(self, slot: int):
        self._ck(self.lib.fdtd_batch_retire(self.b, slot), "fdtd_batch_retire")
        self.members.pop(slot, None)

    def step(self, nsteps: int):
        """nsteps steps of every admitted grid; probe samples are drained into
        each Solver between chunks that fit its probe buffer."""
        left = nsteps
        while left > 0:
            k = left
            for s in self.members.values():
                if s.nprobe:
                    k = min(k, s.probe_buffer)
            for s in self.members.values():
                s._drain_if_full(k)
            self._ck(self.lib.fdtd_batch_step(self.b, k), "fdtd_batch_step")
            for s in self.members.values():
                s.n += k
            left -= k

    def synchronize(self):
        self._ck(self.lib.fdtd_batch_synchronize(self.b), "fdtd_batch_synchronize")

    def close(self):
        if getattr(self, "b", None):
            for slot in list(self.members):
                self.retire(slot)
            self.lib.fdtd_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
