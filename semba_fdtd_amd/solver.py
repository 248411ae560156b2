"""Host-side mirror of semba-fdtd's ``launch_simulation`` for the time-stepping
hot path, driving the gfx950 library through its C ABI.

Set-up follows the reference order (timestepping.F90:305-1014): spacings,
calc_G1G2Gm1Gm2, InitCPMLBorders (which stretches the inverse spacings),
InitEDispersives (which overrides G1/G2 of the dispersive media), nodal
sources, plane waves, probes.  The step methods keep the reference's call structure:
``advance_e(n)`` / ``advance_h(n)`` are the E and H halves of one iteration of
``ciclo_temporal`` (timestepping.F90:1248-1981) and ``run`` replays whole steps
from a captured HIP graph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List
import numpy as np

from . import _lib
from ._lib import FdtdError, f32p, f64p, i16p, i32p, ptr
from .problem import Problem, PMAX, KIND_EDISP, KIND_MDISP, alloc_extent, to_alloc_box

MODES = {"fast": _lib.MODE_FAST, "exact": _lib.MODE_EXACT}
STORAGE = {"fp32": _lib.STORE_FP32, "fp16": _lib.STORE_FP16}


@dataclass
class Tables:
    g1: np.ndarray
    g2: np.ndarray
    gm1: np.ndarray
    gm2: np.ndarray
    ide: List[np.ndarray]
    idh: List[np.ndarray]
    pbe: List[np.ndarray]
    pce: List[np.ndarray]
    pbm: List[np.ndarray]
    pcm: List[np.ndarray]
    disp_media: np.ndarray
    disp_npoles: np.ndarray
    kappa: np.ndarray
    beta: np.ndarray
    g3: np.ndarray
    lines: List[np.ndarray] = None
    mdisp_media: np.ndarray = None
    mdisp_npoles: np.ndarray = None
    mkappa: np.ndarray = None
    mbeta: np.ndarray = None
    gm3: np.ndarray = None
    cab1: np.ndarray = None             # [6][nmedia+1] first-order Mur, or None


def _check(lib, h, rc, what):
    if rc != 0:
        msg = lib.fdtd_last_error(h)
        raise FdtdError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def build_tables(pb: Problem) -> Tables:
    """Every coefficient table of ``pb``, in the reference's set-up order."""
    lib = _lib.load()
    nm = pb.nmedia
    z = lambda n: np.zeros(n, np.float32)
    g1, g2, gm1, gm2 = z(nm + 1), z(nm + 1), z(nm + 1), z(nm + 1)
    rc = lib.fdtd_build_g_tables(nm, ptr(pb.epr, f32p), ptr(pb.sigma, f32p), ptr(pb.mur, f32p),
                                 ptr(pb.sigmam, f32p), ptr(pb.kind, i32p), pb.dt, pb.eps0,
                                 pb.mu0, ptr(g1, f32p), ptr(g2, f32p), ptr(gm1, f32p),
                                 ptr(gm2, f32p))
    if rc:
        raise FdtdError("fdtd_build_g_tables failed")
    ide, idh, pbe, pce, pbm, pcm = [], [], [], [], [], []
    for a, n in enumerate((pb.nx, pb.ny, pb.nz)):
        de, dh, Ie, Ih = z(n + 3), z(n + 3), z(n + 3), z(n + 3)
        lib.fdtd_build_spacing(n, ptr(pb.D[a], f32p), ptr(de, f32p), ptr(dh, f32p),
                               ptr(Ie, f32p), ptr(Ih, f32p))
        b1, c1, b2, c2 = np.ones(n + 3, np.float32), z(n + 3), np.ones(n + 3, np.float32), z(n + 3)
        L = pb.cpml_layers[a]
        # the reference computes the profiles of all three axes as soon as one
        # face has layers (borderscpml.F90:228-300, 1465-1733)
        if pb.has_cpml:
            rc = lib.fdtd_build_cpml_axis(
                n, int(L[0]), int(L[1]), float(pb.cpml_order[a][0]), float(pb.cpml_order[a][1]),
                float(pb.cpml_refl[a][0]), float(pb.cpml_refl[a][1]), pb.alphamax, pb.alphaorden,
                pb.kappamax, pb.dt, pb.eps0, pb.mu0, ptr(de, f32p), ptr(dh, f32p), ptr(b1, f32p),
                ptr(c1, f32p), ptr(b2, f32p), ptr(c2, f32p), ptr(Ie, f32p), ptr(Ih, f32p))
            if rc:
                raise FdtdError(f"fdtd_build_cpml_axis failed on axis {a}")
        ide.append(Ie); idh.append(Ih)
        pbe.append(b1); pce.append(c1); pbm.append(b2); pcm.append(c2)
    disp = [m for m in range(1, nm + 1) if pb.kind[m] & KIND_EDISP]
    nd = len(disp)
    kappa = np.zeros((max(nd, 1), PMAX, 2)); beta = np.zeros_like(kappa); g3 = np.zeros_like(kappa)
    for d, m in enumerate(disp):
        a1, a2 = np.zeros(1, np.float32), np.zeros(1, np.float32)
        c = np.ascontiguousarray(pb.pole_c[m]); a = np.ascontiguousarray(pb.pole_a[m])
        kk = np.zeros((PMAX, 2)); bb = np.zeros((PMAX, 2)); gg = np.zeros((PMAX, 2))
        rc = lib.fdtd_build_ade(int(pb.npoles[m]), ptr(c, f64p), ptr(a, f64p), float(pb.eps11[m]),
                                float(pb.sig11[m]), pb.dt, ptr(kk, f64p), ptr(bb, f64p),
                                ptr(gg, f64p), ptr(a1, f32p), ptr(a2, f32p))
        if rc:
            raise FdtdError(f"fdtd_build_ade failed for medium {m}")
        kappa[d], beta[d], g3[d] = kk, bb, gg
        g1[m], g2[m] = a1[0], a2[0]
    # magnetic media after the electric ones (InitMDispersives follows
    # InitEDispersives, timestepping.F90:927-943)
    mdisp = [m for m in range(1, nm + 1) if pb.kind[m] & KIND_MDISP]
    nmd = len(mdisp)
    mkappa = np.zeros((max(nmd, 1), PMAX, 2)); mbeta = np.zeros_like(mkappa)
    gm3 = np.zeros_like(mkappa)
    for d, m in enumerate(mdisp):
        a1, a2 = np.zeros(1, np.float32), np.zeros(1, np.float32)
        c = np.ascontiguousarray(pb.mpole_c[m]); a = np.ascontiguousarray(pb.mpole_a[m])
        kk = np.zeros((PMAX, 2)); bb = np.zeros((PMAX, 2)); gg = np.zeros((PMAX, 2))
        rc = lib.fdtd_build_ade(int(pb.mnpoles[m]), ptr(c, f64p), ptr(a, f64p),
                                float(pb.mu11[m]), float(pb.sigm11[m]), pb.dt, ptr(kk, f64p),
                                ptr(bb, f64p), ptr(gg, f64p), ptr(a1, f32p), ptr(a2, f32p))
        if rc:
            raise FdtdError(f"fdtd_build_ade failed for magnetic medium {m}")
        mkappa[d], mbeta[d], gm3[d] = kk, bb, gg
        gm1[m], gm2[m] = a1[0], a2[0]
    # first-order Mur after the CPML stretching (InitMURBorders follows
    # InitCPMLBorders, timestepping.F90:635-669)
    cab1 = None
    if any(pb.mur_faces):
        idf = np.array([idh[a][(n + 1) if hi else 0]
                        for a, n in enumerate((pb.nx, pb.ny, pb.nz)) for hi in (0, 1)],
                       np.float32)
        cab1 = np.zeros((6, nm + 1), np.float32)
        rc = lib.fdtd_build_mur(nm, ptr(pb.epr, f32p), ptr(pb.mur, f32p), pb.dt, pb.eps0,
                                pb.mu0, ptr(idf, f32p), ptr(cab1, f32p))
        if rc:
            raise FdtdError("fdtd_build_mur failed")
    lines = []
    for a, n in enumerate((pb.nx, pb.ny, pb.nz)):
        ln = z(n + 3)
        rc = lib.fdtd_build_lines(n, int(pb.cpml_layers[a][0]), int(pb.cpml_layers[a][1]),
                                  ptr(pb.D[a], f32p), ptr(ln, f32p))
        if rc:
            raise FdtdError(f"fdtd_build_lines failed on axis {a}")
        lines.append(ln)
    return Tables(g1, g2, gm1, gm2, ide, idh, pbe, pce, pbm, pcm,
                  np.array(disp, np.int32), np.array([pb.npoles[m] for m in disp], np.int32),
                  kappa, beta, g3, lines, np.array(mdisp, np.int32),
                  np.array([pb.mnpoles[m] for m in mdisp], np.int32), mkappa, mbeta, gm3, cab1)


ETA0 = 376.730313


def fp16_scales(pb: Problem, t: Tables) -> List[float]:
    """Per-array scales for binary16 field storage: the peak field the sources
    can drive, estimated from their amplitudes, mapped to 4096 so that values
    keep a factor 16 of headroom below the binary16 maximum (65504).  H arrays
    take the E scale over the vacuum impedance."""
    peak = 0.0
    for w in pb.planewaves:
        amp = max(float(np.linalg.norm(e)) for e in w.e)
        peak = max(peak, 2.0 * amp * float(np.abs(w.evol).max()))
    dmin = min(float(d.min()) for d in pb.D)
    for src in pb.nodal:
        ev = np.abs(np.asarray(src.evol, np.float64))
        m = abs(src.amp) * float(ev.max())
        if src.comp >= 3:
            m *= ETA0
        elif not src.hard:
            # a soft current integrates G2 * amp * f / d^2 over the pulse
            width = max(1, int((ev > 0.5 * ev.max()).sum()))
            m *= float(np.max(t.g2)) / (dmin * dmin) * width
        peak = max(peak, m)
    e = (peak if peak > 0 else 1.0) / 4096.0
    return [e, e, e, e / ETA0, e / ETA0, e / ETA0]


class Solver:
    """One grid resident in HBM.  ``mode`` is "fast" (FMA, folded fp32 poles)
    or "exact" (bit-identical to the reference loops); ``storage`` "fp16"
    keeps the fields as scaled binary16 (FAST only)."""

    def __init__(self, pb: Problem, mode: str = "fast", device: int = 0, stream=None,
                 tables: Tables | None = None, upload_media: bool = True, probe_buffer: int = 4096,
                 storage: str = "fp32"):
        self.lib = lib = _lib.load()
        self.pb = pb
        self.n3 = (pb.nx, pb.ny, pb.nz)
        self.mode = mode
        self.storage = storage
        self.t = t = tables if tables is not None else build_tables(pb)
        d = _lib.FdtdGridDesc(pb.nx, pb.ny, pb.nz, MODES[mode], STORAGE[storage], device,
                              stream)
        h = C.c_void_p()
        rc = lib.fdtd_create(C.byref(d), C.byref(h))
        if rc:
            raise FdtdError(f"fdtd_create failed ({rc}): {lib.fdtd_last_error(None).decode()}")
        self.h = h
        self._keep = []
        if storage == "fp16":
            for c, sc in enumerate(fp16_scales(pb, t)):
                self._ck(lib.fdtd_set_field_scale(h, c, sc), "fdtd_set_field_scale")
        self._ck(lib.fdtd_set_media(h, pb.nmedia, ptr(t.g1, f32p), ptr(t.g2, f32p),
                                    ptr(t.gm1, f32p), ptr(t.gm2, f32p), ptr(pb.kind, i32p)),
                 "fdtd_set_media")
        if pb.mi is None:
            self._apply_scene()
        elif upload_media:
            for c in range(6):
                m = to_alloc_box(pb.mi[c].astype(np.int16, copy=False), c, self.n3)
                self._ck(lib.fdtd_set_media_index(h, c, ptr(m, i16p)), "fdtd_set_media_index")
        # Idx/y/zh span Alloc(iEx..iEz): one entry fewer than the primal tables
        ids = t.ide + [np.ascontiguousarray(a[:-1]) for a in t.idh]
        self._ck(lib.fdtd_set_spacing(h, *[ptr(a, f32p) for a in ids]), "fdtd_set_spacing")
        if pb.has_cpml:
            lay = np.array([pb.cpml_layers[a][s] for a in range(3) for s in range(2)], np.int32)
            prof = []
            for a in range(3):
                prof += [t.pbe[a], t.pce[a], t.pbm[a], t.pcm[a]]
            self._ck(lib.fdtd_set_cpml(h, ptr(lay, i32p), *[ptr(p, f32p) for p in prof]),
                     "fdtd_set_cpml")
        nd = len(t.disp_media)
        if nd:
            dm = np.ascontiguousarray(t.disp_media, np.int32)
            dn = np.ascontiguousarray(t.disp_npoles, np.int32)
            k = np.ascontiguousarray(t.kappa[:nd]); b = np.ascontiguousarray(t.beta[:nd])
            g = np.ascontiguousarray(t.g3[:nd])
            self._keep += [dm, dn, k, b, g]
            self._ck(lib.fdtd_set_edispersive(h, nd, ptr(dm, i32p), ptr(dn, i32p), ptr(k, f64p),
                                              ptr(b, f64p), ptr(g, f64p)), "fdtd_set_edispersive")
        nmd = len(t.mdisp_media) if t.mdisp_media is not None else 0
        if nmd:
            dm = np.ascontiguousarray(t.mdisp_media, np.int32)
            dn = np.ascontiguousarray(t.mdisp_npoles, np.int32)
            k = np.ascontiguousarray(t.mkappa[:nmd]); b = np.ascontiguousarray(t.mbeta[:nmd])
            g = np.ascontiguousarray(t.gm3[:nmd])
            self._keep += [dm, dn, k, b, g]
            self._ck(lib.fdtd_set_mdispersive(h, nmd, ptr(dm, i32p), ptr(dn, i32p), ptr(k, f64p),
                                              ptr(b, f64p), ptr(g, f64p)), "fdtd_set_mdispersive")
        if t.cab1 is not None:
            faces = np.array(pb.mur_faces, np.int32)
            cab = np.ascontiguousarray(t.cab1, np.float32)
            self._ck(lib.fdtd_set_mur(h, ptr(faces, i32p), ptr(cab, f32p)), "fdtd_set_mur")
        nlast = max(pb.nsteps_max + 2, 2)
        self._ck(lib.fdtd_set_time(h, pb.dt, nlast, None), "fdtd_set_time")
        for s in pb.nodal:
            box = np.array(s.box, np.int32)
            ev = np.ascontiguousarray(s.evol, np.float32)
            self._ck(lib.fdtd_add_nodal(h, s.comp, int(s.hard), ptr(box, i32p), s.amp,
                                        s.deltaevol, len(ev) - 1, ptr(ev, f32p)), "fdtd_add_nodal")
        if pb.planewaves:
            self._ck(lib.fdtd_set_lines(h, *[ptr(t.lines[a], f32p) for a in range(3)]),
                     "fdtd_set_lines")
        for w in pb.planewaves:
            box = np.array(w.box, np.int32)
            pp = np.ascontiguousarray(np.array(w.p, np.float32))
            ee = np.ascontiguousarray(np.array(w.e, np.float32))
            inc = np.array(w.incert, np.float32)
            ev = np.ascontiguousarray(w.evol, np.float32)
            self._keep += [box, pp, ee, inc, ev]
            self._ck(lib.fdtd_add_planewave(h, ptr(box, i32p), len(w.p), ptr(pp, f32p),
                                            ptr(ee, f32p), ptr(inc, f32p), w.deltaevol,
                                            len(ev) - 1, ptr(ev, f32p), pb.eps0, pb.mu0),
                     "fdtd_add_planewave")
        self.nprobe = len(pb.probes)
        self.probe_buffer = probe_buffer
        if self.nprobe:
            P = np.array(pb.probes, np.int32).T.copy()
            self._ck(lib.fdtd_set_probes(h, self.nprobe, ptr(P[0], i32p), ptr(P[1], i32p),
                                         ptr(P[2], i32p), ptr(P[3], i32p), probe_buffer),
                     "fdtd_set_probes")
        self.dft_ids = []
        for d in pb.dfts:
            box = np.array(d.box, np.int32)
            st = np.array(d.stride, np.int32)
            did = C.c_int32()
            self._ck(lib.fdtd_add_dft(h, int(d.kind), ptr(box, i32p), ptr(st, i32p),
                                      len(d.freqs), ptr(d.freqs, f32p), C.byref(did)),
                     "fdtd_add_dft")
            self.dft_ids.append(did.value)
        self.n = 0
        self._samples = []

    def _ck(self, rc, what):
        _check(self.lib, self.h, rc, what)

    def _apply_scene(self):
        """Build the media matrices in HBM from the problem's scene ops."""
        lib, h, pb = self.lib, self.h, self.pb
        prio = np.ascontiguousarray(pb.prio, np.int32)
        npr = len(prio)
        for op in pb.scene:
            kind = op[0]
            if kind == "fill":
                rc = lib.fdtd_media_fill(h, int(op[1]))
            elif kind == "pec":
                rc = lib.fdtd_media_pec_walls(h, int(op[1]))
            elif kind == "sphere":
                rc = lib.fdtd_media_paint_sphere(h, *[float(v) for v in op[1:5]], int(op[5]),
                                                 npr, ptr(prio, i32p))
            elif kind == "box":
                rc = lib.fdtd_media_paint_box(h, *[int(v) for v in op[1:7]], int(op[7]), npr,
                                              ptr(prio, i32p))
            else:
                raise ValueError(f"unknown scene op {kind}")
            self._ck(rc, f"scene op {kind}")

    # -- the reference's call structure ------------------------------------
    def advance_e(self, n: int):
        self._ck(self.lib.fdtd_advance_e(self.h, n), "fdtd_advance_e")

    def advance_h(self, n: int):
        self._ck(self.lib.fdtd_advance_h(self.h, n), "fdtd_advance_h")

    def step(self):
        """One iteration of ciclo_temporal, then UpdateObservation."""
        self.advance_e(self.n)
        self.advance_h(self.n)
        if self.nprobe or self.dft_ids:
            self._drain_if_full(1)
            self._ck(self.lib.fdtd_update_observation(self.h, self.n), "fdtd_update_observation")
        self.n += 1

    def dft_sums(self):
        """Running sums of each DftBox of the problem, complex128 arrays of
        DftBox.shape."""
        out = []
        for d, did in zip(self.pb.dfts, self.dft_ids):
            n = int(self.lib.fdtd_dft_size(self.h, did))
            buf = np.zeros((n, 2))
            self._ck(self.lib.fdtd_read_dft(self.h, did, ptr(buf, f64p), n), "fdtd_read_dft")
            out.append((buf[:, 0] + 1j * buf[:, 1]).reshape(d.shape))
        return out

    def run(self, nsteps: int, chunk: int | None = None):
        """nsteps whole steps from the captured graph; probes are drained from
        the device buffer between chunks."""
        chunk = chunk or self.probe_buffer
        left = nsteps
        while left > 0:
            k = min(left, chunk, self.probe_buffer)
            self._drain_if_full(k)
            self._ck(self.lib.fdtd_run(self.h, self.n, k), "fdtd_run")
            self.n += k
            left -= k

    def _drain_if_full(self, need):
        if self.nprobe and getattr(self, "_pending", 0) + need > self.probe_buffer:
            self._drain()
        self._pending = getattr(self, "_pending", 0) + need

    def _drain(self):
        if not self.nprobe:
            return
        out = np.zeros((self.probe_buffer, self.nprobe), np.float32)
        ns = C.c_int32()
        self._ck(self.lib.fdtd_read_probes(self.h, ptr(out, f32p), self.probe_buffer, C.byref(ns)),
                 "fdtd_read_probes")
        self._samples.append(out[:ns.value].copy())
        self._pending = 0

    def probes(self) -> np.ndarray:
        """All probe samples so far, [step][probe]."""
        self._drain()
        if not self._samples:
            return np.zeros((0, self.nprobe), np.float32)
        return np.concatenate(self._samples, axis=0)

    def synchronize(self):
        self._ck(self.lib.fdtd_synchronize(self.h), "fdtd_synchronize")

    def field(self, c: int) -> np.ndarray:
        """Component c on the uniform (nz+3, ny+3, nx+3) box; positions
        outside the component's own Alloc box read as zero."""
        ex, ey, ez = alloc_extent(c, self.n3)
        box = np.zeros((ez, ey, ex), np.float32)
        self._ck(self.lib.fdtd_get_field(self.h, c, ptr(box, f32p)), "fdtd_get_field")
        out = np.zeros(self.pb.alloc_shape, np.float32)
        out[:ez, :ey, :ex] = box
        return out

    def set_field(self, c: int, a: np.ndarray):
        """Load component c from a uniform (nz+3, ny+3, nx+3) array."""
        box = to_alloc_box(np.asarray(a, np.float32), c, self.n3)
        self._ck(self.lib.fdtd_set_field(self.h, c, ptr(box, f32p)), "fdtd_set_field")

    def save_state(self) -> bytes:
        """The whole time-stepping state at the current step (a checkpoint)."""
        self._drain()
        size = int(self.lib.fdtd_state_size(self.h))
        if size < 0:
            raise FdtdError(f"fdtd_state_size failed: {self.lib.fdtd_last_error(self.h).decode()}")
        buf = C.create_string_buffer(size)
        self._ck(self.lib.fdtd_save_state(self.h, self.n, buf, size), "fdtd_save_state")
        return buf.raw

    def load_state(self, state: bytes):
        """Resume from a checkpoint of a grid with the same set-up and mode."""
        n = C.c_int32()
        buf = C.create_string_buffer(state, len(state))
        self._ck(self.lib.fdtd_load_state(self.h, buf, len(state), C.byref(n)), "fdtd_load_state")
        self.n = n.value

    def set_tiling(self, rows: int, zchunk: int, xcells: int = 1):
        """Blocks of 64 x rows threads marching zchunk z planes, xcells x
        cells per lane of the fused passes."""
        self._ck(self.lib.fdtd_set_tiling(self.h, rows, zchunk, xcells), "fdtd_set_tiling")

    def get_tiling(self):
        r, k, x = C.c_int32(), C.c_int32(), C.c_int32()
        self._ck(self.lib.fdtd_get_tiling(self.h, C.byref(r), C.byref(k), C.byref(x)),
                 "fdtd_get_tiling")
        return r.value, k.value, x.value

    def set_graph_steps(self, steps: int):
        """Two-pass steps recorded per graph replay in run() (1..64)."""
        self._ck(self.lib.fdtd_set_graph_steps(self.h, steps), "fdtd_set_graph_steps")

    def set_step_mode(self, rows: int, zchunk: int = 32):
        """One launch per step in run() (rows 4, 8 or 16 per block, 0: two
        passes per step).  Needs fp32 storage; holds a second copy of the
        time-stepping state."""
        self._ck(self.lib.fdtd_set_step_mode(self.h, rows, zchunk), "fdtd_set_step_mode")

    STEP_CLASSES = 17

    def set_step_class_rows(self, rows: int, classes: int = 0x1FFFF):
        """Rows per block (4, 6, 8 or 16; 0: the default) of the step kernels
        of the tile classes in the bit mask ``classes`` (bit c: class c; for
        c < 8 bits 1 CPML, 2 poles, 4 sources; class 8 the plain tiles run
        four x cells per lane, 4 or 8 rows; classes 9-16 the half-width
        variants of classes 0-7, in waves per block).  A class on fewer rows
        than the step runs the step's tiles cut into row bands.  The results
        do not change."""
        self._ck(self.lib.fdtd_set_step_class_rows(self.h, classes, rows),
                 "fdtd_set_step_class_rows")

    def set_step_vector(self, on: bool):
        """Plain tiles of the one-launch step four x cells per lane (the
        default) or one."""
        self._ck(self.lib.fdtd_set_step_vector(self.h, 1 if on else 0), "fdtd_set_step_vector")

    def get_step_class_rows(self):
        """Rows per block each of the tile classes runs with."""
        r = (C.c_int32 * self.STEP_CLASSES)()
        self._ck(self.lib.fdtd_get_step_class_rows(self.h, r), "fdtd_get_step_class_rows")
        return [int(v) for v in r]

    def step_classes(self, reps: int = 0):
        """The tile classes of the one-launch step: [(class, owned Yee cells,
        ms per launch or None)]: bits 1 CPML, 2 poles, 4 sources; 8 the plain
        tiles four x cells per lane; 9 + bits the half-width tiles."""
        cls = (C.c_int32 * self.STEP_CLASSES)()
        cells = (C.c_double * self.STEP_CLASSES)()
        nbytes = (C.c_double * self.STEP_CLASSES)()
        ms = (C.c_float * self.STEP_CLASSES)()
        n = self.lib.fdtd_step_classes(self.h, reps, cls, cells, nbytes, ms)
        self._ck(min(n, 0), "fdtd_step_classes")
        self._class_bytes = {int(cls[q]): float(nbytes[q]) for q in range(n)}
        return [(int(cls[q]), float(cells[q]), float(ms[q]) if reps else None) for q in range(n)]

    def step_class_bytes(self):
        """{class: compulsory HBM bytes of one launch of its kernel} (fields,
        media, CPML psi and pole state of the cells its tiles own)."""
        self.step_classes()
        return dict(self._class_bytes)

    def get_step_mode(self):
        r, k = C.c_int32(), C.c_int32()
        self._ck(self.lib.fdtd_get_step_mode(self.h, C.byref(r), C.byref(k)), "fdtd_get_step_mode")
        return r.value, k.value

    def time_step(self, reps: int = 3) -> float:
        """Average ms of one one-launch step.  Advances the fields."""
        ms = np.zeros(1, np.float32)
        self._ck(self.lib.fdtd_time_step(self.h, reps, ptr(ms, f32p)), "fdtd_time_step")
        return float(ms[0])

    def time_halves(self, reps: int = 5):
        """Average ms of the fused E pass (with poles), the fused H pass and
        the pole update alone.  Advances the fields."""
        ms = np.zeros(3, np.float32)
        self._ck(self.lib.fdtd_time_halves(self.h, reps, ptr(ms, f32p)), "fdtd_time_halves")
        return tuple(float(v) for v in ms)

    @property
    def device_bytes(self) -> int:
        return int(self.lib.fdtd_device_bytes(self.h))

    def close(self):
        """Free the grid (refused while it is admitted to a batch)."""
        if getattr(self, "h", None):
            self._ck(self.lib.fdtd_destroy(self.h), "fdtd_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
