"""CPU oracle for semba-fdtd's field time-stepping path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product in
``semba_fdtd_amd`` never does.

Two checkers live here:

* ``Oracle`` -- ctypes wrapper of ``fdtd_oracle.c``, a plain-C restatement of
  the reference's loops and coefficient tables (no FMA contraction), built in
  the reference's real(4) kind (``liboracle_fdtd.so``) and in real(8)
  (``liboracle_fdtd64.so``, the FP64 yardstick, ``precision=64``).
* ``Reference`` -- ctypes wrapper of ``_ref/libsemba_ref.so``: the reference's
  own Fortran modules (calc_constants, borderscpml, electricdispersive,
  magneticdispersive, bordersmur, nodalsources, planewaves) compiled from
  their sources by ``oracle/Makefile``, driven by ``ref_harness.F90``.  Only
  built where the reference tree is present.

Both consume a ``semba_fdtd_amd.problem.Problem`` (plain numpy data: media
tables, int16 media-index arrays in the Alloc layout, cell sizes, CPML
settings, poles, nodal sources, probes).
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = {32: os.path.join(HERE, "liboracle_fdtd.so"),
             64: os.path.join(HERE, "liboracle_fdtd64.so")}
# the real(4) build with FMA contraction allowed (x86-64-v3): only for timing
# the CPU baseline with FMA on; parity uses the builds above
ORACLE_FMA_SO = os.path.join(HERE, "liboracle_fdtd_fma.so")
REF_SO = os.path.join(HERE, "_ref", "libsemba_ref.so")
# the same modules built with OpenMP: only for timing the reference on the
# host's cores
REF_OMP_SO = os.path.join(HERE, "_ref_omp", "libsemba_ref.so")
PMAX = 8
PW_MODES = 4
# vacuum constants as written in the reference (fdetypes.F90:219-222), kept
# in full for the real(8) build
EPS0_64 = 8.8541878176203898505365630317107502606083701665994498081024171524053950954599821142852891607182008932e-12
MU0_64 = 1.2566370614359172953850573533118011536788677597500423283899778369231265625144835994512139301368468271e-6

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i16p = C.POINTER(C.c_int16)
_REAL = {32: (C.c_float, np.float32), 64: (C.c_double, np.float64)}


def _structs(prec):
    rt = _REAL[prec][0]
    rp = C.POINTER(rt)

    class OrcNodal(C.Structure):
        _fields_ = [("comp", C.c_int32), ("hard", C.c_int32),
                    ("xi", C.c_int32), ("xe", C.c_int32), ("yi", C.c_int32),
                    ("ye", C.c_int32), ("zi", C.c_int32), ("ze", C.c_int32),
                    ("amp", rt), ("deltaevol", rt), ("numus", C.c_int32), ("evol", rp)]

    class OrcPW(C.Structure):
        _fields_ = [("esq", C.c_int32 * 6), ("nmodes", C.c_int32),
                    ("p", (rt * 3) * PW_MODES), ("e", (rt * 3) * PW_MODES),
                    ("incert", rt * PW_MODES), ("deltaevol", rt), ("numus", C.c_int32),
                    ("evol", rp)]

    class OrcGrid(C.Structure):
        _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                    ("f", rp * 6), ("mi", _i16p * 6),
                    ("g1", rp), ("g2", rp), ("gm1", rp), ("gm2", rp),
                    ("nmedia", C.c_int32), ("id", rp * 6),
                    ("L", (C.c_int32 * 2) * 3),
                    ("pbe", rp * 3), ("pce", rp * 3), ("pbm", rp * 3), ("pcm", rp * 3),
                    ("psi", rp * 12),
                    ("ndisp", C.c_int32), ("disp_medium", _i32p),
                    ("disp_npoles", _i32p), ("disp_kappa", _f64p),
                    ("disp_beta", _f64p), ("disp_g3", _f64p),
                    ("eprev", rp * 3), ("jcur", _f64p * 3), ("jstride", C.c_int32),
                    ("dt", C.c_double), ("tiempo", _f64p),
                    ("nnodal", C.c_int32), ("nodal", C.POINTER(OrcNodal)),
                    ("medkind", _i32p),
                    ("npw", C.c_int32), ("pw", C.POINTER(OrcPW)), ("line", rp * 3),
                    ("eps0", rt), ("mu0", rt), ("pws", C.c_void_p),
                    ("nmdisp", C.c_int32), ("mdisp_medium", _i32p),
                    ("mdisp_npoles", _i32p), ("mdisp_kappa", _f64p),
                    ("mdisp_beta", _f64p), ("mdisp_gm3", _f64p),
                    ("hprev", rp * 3), ("mjcur", _f64p * 3),
                    ("mur", C.c_int32 * 6), ("cab1", rp * 6), ("murpast", rp * 3)]
    return OrcNodal, OrcGrid, OrcPW


_STRUCTS = {p: _structs(p) for p in (32, 64)}
OrcNodal, OrcGrid, OrcPW = _STRUCTS[32]


def _p(a, t):
    return a.ctypes.data_as(t)


_LIBS = {}


def load_oracle(precision: int = 32, fma: bool = False):
    key = (precision, fma)
    if key in _LIBS:
        return _LIBS[key]
    if fma and precision != 32:
        raise ValueError("the FMA build is real(4) only")
    path = ORACLE_FMA_SO if fma else ORACLE_SO[precision]
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
    rt = _REAL[precision][0]
    rp = C.POINTER(rt)
    Nod, Grid, _ = _STRUCTS[precision]
    lib = C.CDLL(path)
    lib.orc_dt.restype = C.c_double
    lib.orc_dt.argtypes = [rt] * 6 + [C.c_int32, C.c_double]
    lib.orc_g_tables.argtypes = [C.c_int32, rp, rp, rp, rp, _i32p, C.c_double, rt, rt,
                                 rp, rp, rp, rp]
    lib.orc_spacing.argtypes = [C.c_int32, rp, rp, rp, rp, rp]
    lib.orc_cpml_axis.argtypes = ([C.c_int32] * 3 + [rt] * 7 + [C.c_double, rt, rt] +
                                  [rp] * 8)
    lib.orc_ade_tables.argtypes = [C.c_int32, _f64p, _f64p, rt, rt, C.c_double,
                                   _f64p, _f64p, _f64p, rp, rp]
    lib.orc_time_vector.argtypes = [C.c_int32, C.c_double, _f64p]
    lib.orc_mur_table.argtypes = [C.c_int32, rp, rp, C.c_double, rt, rt, rt, rp]
    lib.orc_evolucion.restype = rt
    lib.orc_evolucion.argtypes = [rt, C.POINTER(Nod)]
    for fn in ("orc_yee_e", "orc_yee_h", "orc_cpml_e", "orc_cpml_h", "orc_ade_e", "orc_ade_h",
               "orc_mur_h"):
        getattr(lib, fn).argtypes = [C.POINTER(Grid)]
    for fn in ("orc_nodal_e", "orc_nodal_h", "orc_step", "orc_pw_e", "orc_pw_h"):
        getattr(lib, fn).argtypes = [C.POINTER(Grid), C.c_int32]
    lib.orc_pw_init.argtypes = [C.POINTER(Grid)]
    lib.orc_pw_init.restype = C.c_int
    lib.orc_pw_free.argtypes = [C.POINTER(Grid)]
    lib.orc_lines.argtypes = [C.c_int32, C.c_int32, C.c_int32, rp, rp]
    lib.orc_dft_aux.argtypes = [C.c_int32, _f32p, C.c_double, C.c_double, C.c_int32, _f64p]
    lib.orc_dft_accum.argtypes = [C.POINTER(Grid), C.c_int32, _i32p, _i32p, C.c_int32, _f64p,
                                  _f64p]
    lib.orc_run.argtypes = [C.POINTER(Grid), C.c_int32, C.c_int32, C.c_int32,
                            _i32p, _i32p, _i32p, _i32p, rp]
    _LIBS[key] = lib
    return lib


class Tables:
    """Coefficient tables of one problem (as the reference builds them)."""

    def __init__(self, g1, g2, gm1, gm2, ide, idh, pbe, pce, pbm, pcm,
                 kappa, beta, g3, disp_media, disp_npoles, mkappa=None, mbeta=None,
                 gm3=None, mdisp_media=None, mdisp_npoles=None, cab1=None):
        self.g1, self.g2, self.gm1, self.gm2 = g1, g2, gm1, gm2
        self.ide, self.idh = ide, idh          # lists of 3 arrays (x,y,z)
        self.pbe, self.pce, self.pbm, self.pcm = pbe, pce, pbm, pcm
        self.kappa, self.beta, self.g3 = kappa, beta, g3
        self.disp_media, self.disp_npoles = disp_media, disp_npoles
        z = np.zeros((1, PMAX, 2))
        self.mkappa = z if mkappa is None else mkappa
        self.mbeta = z if mbeta is None else mbeta
        self.gm3 = z if gm3 is None else gm3
        self.mdisp_media = np.zeros(0, np.int32) if mdisp_media is None else mdisp_media
        self.mdisp_npoles = np.zeros(0, np.int32) if mdisp_npoles is None else mdisp_npoles
        self.cab1 = cab1                       # [6][nmedia+1] or None


def _constants(pb, precision):
    if precision == 64:
        return EPS0_64, MU0_64
    return pb.eps0, pb.mu0


def oracle_tables(lib, pb, precision: int = 32) -> Tables:
    """Build every coefficient table with the oracle's restatement."""
    rt, dt_ = _REAL[precision]
    rp = C.POINTER(rt)
    nm = pb.nmedia
    f = lambda n: np.zeros(n, dt_)
    r = lambda a: np.ascontiguousarray(a, dt_)
    eps0, mu0 = _constants(pb, precision)
    g1, g2, gm1, gm2 = f(nm + 1), f(nm + 1), f(nm + 1), f(nm + 1)
    epr, sig, mur, sgm = r(pb.epr), r(pb.sigma), r(pb.mur), r(pb.sigmam)
    lib.orc_g_tables(nm, _p(epr, rp), _p(sig, rp), _p(mur, rp), _p(sgm, rp),
                     _p(pb.kind, _i32p), pb.dt, eps0, mu0,
                     _p(g1, rp), _p(g2, rp), _p(gm1, rp), _p(gm2, rp))
    ide, idh, pbe, pce, pbm, pcm = [], [], [], [], [], []
    for a, n in enumerate((pb.nx, pb.ny, pb.nz)):
        de, dh, Ie, Ih = f(n + 3), f(n + 3), f(n + 3), f(n + 3)
        D = r(pb.D[a])
        lib.orc_spacing(n, _p(D, rp), _p(de, rp), _p(dh, rp), _p(Ie, rp), _p(Ih, rp))
        b1, c1, b2, c2 = f(n + 3), f(n + 3), f(n + 3), f(n + 3)
        L = pb.cpml_layers[a]
        # InitCPMLBorders computes the profiles of all three axes as soon as
        # one face has layers (borderscpml.F90:228-300, 1465-1733)
        if pb.has_cpml:
            lib.orc_cpml_axis(n, int(L[0]), int(L[1]), float(pb.cpml_order[a][0]),
                              float(pb.cpml_order[a][1]), float(pb.cpml_refl[a][0]),
                              float(pb.cpml_refl[a][1]), pb.alphamax, pb.alphaorden,
                              pb.kappamax, pb.dt, eps0, mu0,
                              _p(de, rp), _p(dh, rp), _p(b1, rp), _p(c1, rp),
                              _p(b2, rp), _p(c2, rp), _p(Ie, rp), _p(Ih, rp))
        else:
            b1[:] = 1.0
            b2[:] = 1.0
        ide.append(Ie); idh.append(Ih)
        pbe.append(b1); pce.append(c1); pbm.append(b2); pcm.append(c2)
    disp = [m for m in range(1, nm + 1) if pb.kind[m] & 4]
    nd = len(disp)
    kappa = np.zeros((max(nd, 1), PMAX, 2)); beta = np.zeros_like(kappa); g3 = np.zeros_like(kappa)
    for d, m in enumerate(disp):
        a1, a2 = f(1), f(1)
        np_ = int(pb.npoles[m])
        cc = np.ascontiguousarray(pb.pole_c[m]); aa = np.ascontiguousarray(pb.pole_a[m])
        kk = np.zeros((PMAX, 2)); bb = np.zeros((PMAX, 2)); gg = np.zeros((PMAX, 2))
        eps11 = float(pb.eps11[m]) if precision == 32 else float(pb.eps11[m]) / pb.eps0 * EPS0_64
        lib.orc_ade_tables(np_, _p(cc, _f64p), _p(aa, _f64p), eps11,
                           float(pb.sig11[m]), pb.dt, _p(kk, _f64p), _p(bb, _f64p),
                           _p(gg, _f64p), _p(a1, rp), _p(a2, rp))
        kappa[d], beta[d], g3[d] = kk, bb, gg
        g1[m], g2[m] = a1[0], a2[0]
    # magnetic media after the electric ones (InitMDispersives follows
    # InitEDispersives, timestepping.F90:927-943)
    mdisp = [m for m in range(1, nm + 1) if pb.kind[m] & 8]
    nmd = len(mdisp)
    mkappa = np.zeros((max(nmd, 1), PMAX, 2)); mbeta = np.zeros_like(mkappa)
    gm3 = np.zeros_like(mkappa)
    for d, m in enumerate(mdisp):
        a1, a2 = f(1), f(1)
        cc = np.ascontiguousarray(pb.mpole_c[m]); aa = np.ascontiguousarray(pb.mpole_a[m])
        kk = np.zeros((PMAX, 2)); bb = np.zeros((PMAX, 2)); gg = np.zeros((PMAX, 2))
        mu11 = float(pb.mu11[m]) if precision == 32 else float(pb.mu11[m]) / pb.mu0 * MU0_64
        lib.orc_ade_tables(int(pb.mnpoles[m]), _p(cc, _f64p), _p(aa, _f64p), mu11,
                           float(pb.sigm11[m]), pb.dt, _p(kk, _f64p), _p(bb, _f64p),
                           _p(gg, _f64p), _p(a1, rp), _p(a2, rp))
        mkappa[d], mbeta[d], gm3[d] = kk, bb, gg
        gm1[m], gm2[m] = a1[0], a2[0]
    # first-order Mur (InitMURBorders after InitCPMLBorders, :635-669): the
    # dual inverse spacings at the Alloc ends, after any CPML stretching
    cab1 = None
    if any(pb.mur_faces):
        cab1 = np.zeros((6, nm + 1), dt_)
        for fce in range(6):
            a, hi = fce // 2, fce & 1
            n = (pb.nx, pb.ny, pb.nz)[a]
            idh_ = idh[a][n + 1] if hi else idh[a][0]
            lib.orc_mur_table(nm, _p(epr, rp), _p(mur, rp), pb.dt, eps0, mu0, idh_,
                              _p(cab1[fce], rp))
    return Tables(g1, g2, gm1, gm2, ide, idh, pbe, pce, pbm, pcm, kappa, beta, g3,
                  np.array(disp, np.int32), np.array([pb.npoles[m] for m in disp], np.int32),
                  mkappa, mbeta, gm3, np.array(mdisp, np.int32),
                  np.array([pb.mnpoles[m] for m in mdisp], np.int32), cab1)


def lines(lib, pb, precision: int = 32):
    """Node coordinates LineX/Y/Z over [-1, n+1] (orc_lines)."""
    rt, dt_ = _REAL[precision]
    out = []
    for a, n in enumerate((pb.nx, pb.ny, pb.nz)):
        D = np.ascontiguousarray(pb.D[a], dt_)
        ln = np.zeros(n + 3, dt_)
        lib.orc_lines(n, int(pb.cpml_layers[a][0]), int(pb.cpml_layers[a][1]),
                      _p(D, C.POINTER(rt)), _p(ln, C.POINTER(rt)))
        out.append(ln)
    return out


class Oracle:
    """Runs a Problem through the C restatement (precision 32 or 64)."""

    def __init__(self, pb, tables: Tables | None = None, precision: int = 32, fma: bool = False):
        self.precision = precision
        self.lib = load_oracle(precision, fma)
        rt, dt_ = _REAL[precision]
        rp = C.POINTER(rt)
        Nod, Grid, PW = _STRUCTS[precision]
        self.pb = pb
        self.t = tables if tables is not None else oracle_tables(self.lib, pb, precision)
        t = self.t
        if pb.mi is None:
            pb.mi = pb.materialize()
        shape = pb.alloc_shape
        self.fields = np.zeros((6,) + shape, dt_)
        self.mi = np.ascontiguousarray(pb.mi, dtype=np.int16)
        self.psi = np.zeros((12,) + shape, dt_) if pb.has_cpml else None
        nd = len(t.disp_media)
        nmd = len(t.mdisp_media)
        self.eprev = np.zeros((3,) + shape, dt_) if nd else None
        self.hprev = np.zeros((3,) + shape, dt_) if nmd else None
        self.jstride = int(max(list(t.disp_npoles) + list(t.mdisp_npoles) + [1]))
        self.jcur = np.zeros((3,) + shape + (self.jstride, 2)) if nd else None
        self.mjcur = np.zeros((3,) + shape + (self.jstride, 2)) if nmd else None
        self.murpast = np.zeros((3,) + shape, dt_) if t.cab1 is not None else None
        self.tiempo = np.zeros(pb.nsteps_max + 3)
        self.lib.orc_time_vector(pb.nsteps_max + 2, pb.dt, _p(self.tiempo, _f64p))
        self.medkind = np.ascontiguousarray(pb.kind, np.int32)
        self._keep = []
        g = Grid()
        g.nx, g.ny, g.nz = pb.nx, pb.ny, pb.nz
        for c in range(6):
            g.f[c] = self.fields[c].ctypes.data_as(rp)
            g.mi[c] = self.mi[c].ctypes.data_as(_i16p)
        g.g1, g.g2, g.gm1, g.gm2 = (_p(t.g1, rp), _p(t.g2, rp), _p(t.gm1, rp), _p(t.gm2, rp))
        g.nmedia = pb.nmedia
        ids = t.ide + t.idh
        for c in range(6):
            g.id[c] = ids[c].ctypes.data_as(rp)
        for a in range(3):
            for s in range(2):
                g.L[a][s] = int(pb.cpml_layers[a][s])
            g.pbe[a] = _p(t.pbe[a], rp); g.pce[a] = _p(t.pce[a], rp)
            g.pbm[a] = _p(t.pbm[a], rp); g.pcm[a] = _p(t.pcm[a], rp)
        if self.psi is not None:
            for s in range(12):
                g.psi[s] = self.psi[s].ctypes.data_as(rp)
        g.ndisp = nd
        if nd:
            self._dm = np.ascontiguousarray(t.disp_media, np.int32)
            self._dn = np.ascontiguousarray(t.disp_npoles, np.int32)
            self._k = np.ascontiguousarray(t.kappa); self._b = np.ascontiguousarray(t.beta)
            self._g = np.ascontiguousarray(t.g3)
            g.disp_medium = _p(self._dm, _i32p); g.disp_npoles = _p(self._dn, _i32p)
            g.disp_kappa = _p(self._k, _f64p); g.disp_beta = _p(self._b, _f64p)
            g.disp_g3 = _p(self._g, _f64p)
            for c in range(3):
                g.eprev[c] = self.eprev[c].ctypes.data_as(rp)
                g.jcur[c] = self.jcur[c].ctypes.data_as(_f64p)
        g.nmdisp = nmd
        if nmd:
            self._mdm = np.ascontiguousarray(t.mdisp_media, np.int32)
            self._mdn = np.ascontiguousarray(t.mdisp_npoles, np.int32)
            self._mk = np.ascontiguousarray(t.mkappa); self._mb = np.ascontiguousarray(t.mbeta)
            self._mg = np.ascontiguousarray(t.gm3)
            g.mdisp_medium = _p(self._mdm, _i32p); g.mdisp_npoles = _p(self._mdn, _i32p)
            g.mdisp_kappa = _p(self._mk, _f64p); g.mdisp_beta = _p(self._mb, _f64p)
            g.mdisp_gm3 = _p(self._mg, _f64p)
            for c in range(3):
                g.hprev[c] = self.hprev[c].ctypes.data_as(rp)
                g.mjcur[c] = self.mjcur[c].ctypes.data_as(_f64p)
        if t.cab1 is not None:
            self._cab1 = np.ascontiguousarray(t.cab1)
            for fce in range(6):
                g.mur[fce] = int(bool(pb.mur_faces[fce]))
                g.cab1[fce] = self._cab1[fce].ctypes.data_as(rp)
            for c in range(3):
                g.murpast[c] = self.murpast[c].ctypes.data_as(rp)
        g.jstride = self.jstride
        g.dt = pb.dt
        g.tiempo = _p(self.tiempo, _f64p)
        ns = len(pb.nodal)
        self._nod = (Nod * max(ns, 1))()
        for s, src in enumerate(pb.nodal):
            ev = np.ascontiguousarray(src.evol, dt_)
            self._keep.append(ev)
            o = self._nod[s]
            o.comp, o.hard = src.comp, int(src.hard)
            o.xi, o.xe, o.yi, o.ye, o.zi, o.ze = src.box
            o.amp, o.deltaevol, o.numus = src.amp, src.deltaevol, len(ev) - 1
            o.evol = _p(ev, rp)
        g.nnodal = ns
        g.nodal = C.cast(self._nod, C.POINTER(Nod))
        g.medkind = _p(self.medkind, _i32p)
        eps0, mu0 = _constants(pb, precision)
        g.eps0, g.mu0 = eps0, mu0
        self.lines = lines(self.lib, pb, precision)
        for a in range(3):
            g.line[a] = self.lines[a].ctypes.data_as(rp)
        npw = len(pb.planewaves)
        self._pw = (PW * max(npw, 1))()
        for q, w in enumerate(pb.planewaves):
            o = self._pw[q]
            for a in range(6):
                o.esq[a] = int(w.box[a])
            o.nmodes = len(w.p)
            for m in range(len(w.p)):
                for a in range(3):
                    o.p[m][a] = w.p[m][a]
                    o.e[m][a] = w.e[m][a]
                o.incert[m] = w.incert[m]
            ev = np.ascontiguousarray(w.evol, dt_)
            self._keep.append(ev)
            o.deltaevol, o.numus, o.evol = w.deltaevol, len(ev) - 1, _p(ev, rp)
        g.npw = npw
        g.pw = C.cast(self._pw, C.POINTER(PW))
        g.pws = None
        self.g = g
        if npw and self.lib.orc_pw_init(C.byref(g)) != 0:
            raise MemoryError("orc_pw_init failed")
        self.dft = [np.zeros(d.shape + (2,)) for d in pb.dfts]
        self.n = 0

    def _accumulate_dft(self, n):
        for d, acc in zip(self.pb.dfts, self.dft):
            aux = np.zeros((len(d.freqs), 2))
            self.lib.orc_dft_aux(len(d.freqs), _p(d.freqs, _f32p), self.pb.dt, self.tiempo[n],
                                 int(d.kind), _p(aux, _f64p))
            box = np.array(d.box, np.int32)
            st = np.array(d.stride, np.int32)
            self.lib.orc_dft_accum(C.byref(self.g), int(d.kind), _p(box, _i32p), _p(st, _i32p),
                                   len(d.freqs), _p(aux, _f64p), _p(acc, _f64p))

    def dft_sums(self):
        """Running sums of each DftBox as complex128 arrays of DftBox.shape."""
        return [a[..., 0] + 1j * a[..., 1] for a in self.dft]

    def step(self):
        self.lib.orc_step(C.byref(self.g), self.n)
        if self.pb.dfts:
            self._accumulate_dft(self.n)
        self.n += 1

    def __del__(self):
        try:
            if getattr(self, "g", None) is not None and self.g.pws:
                self.lib.orc_pw_free(C.byref(self.g))
        except Exception:
            pass

    def run(self, nsteps, probes=None):
        probes = self.pb.probes if probes is None else probes
        if self.pb.dfts:
            # UpdateObservation after every step: probes, then the DFT sums
            out = [self._run(1, probes) for _ in range(nsteps)]
            return np.concatenate(out, axis=0) if out else self._run(0, probes)
        return self._run(nsteps, probes)

    def _run(self, nsteps, probes):
        npb = len(probes)
        pc = np.array([p[0] for p in probes] or [0], np.int32)
        pi = np.array([p[1] for p in probes] or [0], np.int32)
        pj = np.array([p[2] for p in probes] or [0], np.int32)
        pk = np.array([p[3] for p in probes] or [0], np.int32)
        out = np.zeros((nsteps, max(npb, 1)), _REAL[self.precision][1])
        self.lib.orc_run(C.byref(self.g), self.n, nsteps, npb, _p(pc, _i32p), _p(pi, _i32p),
                         _p(pj, _i32p), _p(pk, _i32p),
                         _p(out, C.POINTER(_REAL[self.precision][0])))
        for s in range(nsteps):
            if self.pb.dfts:
                self._accumulate_dft(self.n + s)
        self.n += nsteps
        return out[:, :npb]


def load_reference(omp: bool = False):
    path = REF_OMP_SO if omp else REF_SO
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.ref_init.argtypes = ([C.c_int32] * 4 + [_f32p] * 4 + [_i32p, _i32p, _f64p, _f64p,
                             _f32p, _f32p, _i32p, _f32p, _f32p] + [C.c_float] * 3 +
                             [C.c_double, _f32p, _f32p, _f32p, _i16p, C.c_float, C.c_float])
    lib.ref_nodal.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _f32p, _f32p, _i32p, _f32p,
                              C.c_int32]
    lib.ref_step.argtypes = [C.c_int32]
    lib.ref_time.argtypes = [C.c_int32]
    lib.ref_planewave.argtypes = [_i32p, C.c_int32, _f32p, _f32p, _f32p, C.c_float, C.c_int32,
                                  _f32p, _f32p, _f32p, _f32p]
    lib.ref_get_field.argtypes = [C.c_int32, _f32p]
    lib.ref_mdisp.argtypes = [C.c_int32, _i32p, _f64p, _f64p, _f32p, _f32p]
    lib.ref_mur.argtypes = [_i32p]
    lib.ref_get_tables.argtypes = [_f32p] * 10
    return lib


class Reference:
    """Runs a Problem through the reference's own Fortran modules.

    The Fortran modules keep their state in module variables, so one
    Reference can exist per process; tests run it in a subprocess."""

    def __init__(self, pb, omp: bool = False):
        lib = load_reference(omp)
        if lib is None:
            raise RuntimeError("reference harness not built (make -C oracle ref)")
        self.lib, self.pb = lib, pb
        if pb.mi is None:
            pb.mi = pb.materialize()
        nm = pb.nmedia
        pc = np.zeros((nm + 1, PMAX, 2)); pa = np.zeros((nm + 1, PMAX, 2))
        for m in range(nm + 1):
            if pb.kind[m] & 4:
                pc[m, :pb.npoles[m]] = pb.pole_c[m][:pb.npoles[m]]
                pa[m, :pb.npoles[m]] = pb.pole_a[m][:pb.npoles[m]]
        lay = np.ascontiguousarray(np.array(pb.cpml_layers, np.int32))
        orden = np.ascontiguousarray(np.array(pb.cpml_order, np.float32))
        refl = np.ascontiguousarray(np.array(pb.cpml_refl, np.float32))
        self._keep = [pc, pa, lay, orden, refl]
        mi = np.ascontiguousarray(pb.mi, np.int16)
        self._keep.append(mi)
        kind = np.ascontiguousarray(pb.kind, np.int32)
        npol = np.ascontiguousarray(pb.npoles, np.int32)
        self._keep += [kind, npol]
        lib.ref_init(pb.nx, pb.ny, pb.nz, nm, _p(pb.epr, _f32p), _p(pb.sigma, _f32p),
                     _p(pb.mur, _f32p), _p(pb.sigmam, _f32p), _p(kind, _i32p),
                     _p(npol, _i32p), _p(pc, _f64p), _p(pa, _f64p),
                     _p(pb.eps11, _f32p), _p(pb.sig11, _f32p), _p(lay, _i32p),
                     _p(orden, _f32p), _p(refl, _f32p), pb.alphamax, pb.alphaorden,
                     pb.kappamax, pb.dt, _p(pb.D[0], _f32p), _p(pb.D[1], _f32p),
                     _p(pb.D[2], _f32p), _p(mi, _i16p), pb.eps0, pb.mu0)
        if any(pb.kind[m] & 8 for m in range(nm + 1)):
            mpc = np.zeros((nm + 1, PMAX, 2)); mpa = np.zeros((nm + 1, PMAX, 2))
            for m in range(nm + 1):
                if pb.kind[m] & 8:
                    k = int(pb.mnpoles[m])
                    mpc[m, :k] = pb.mpole_c[m][:k]
                    mpa[m, :k] = pb.mpole_a[m][:k]
            mnp = np.ascontiguousarray(pb.mnpoles, np.int32)
            self._keep += [mpc, mpa, mnp]
            lib.ref_mdisp(nm, _p(mnp, _i32p), _p(mpc, _f64p), _p(mpa, _f64p),
                          _p(pb.mu11, _f32p), _p(pb.sigm11, _f32p))
        if any(pb.mur_faces):
            faces = np.array(pb.mur_faces, np.int32)
            self._keep.append(faces)
            lib.ref_mur(_p(faces, _i32p))
        ns = len(pb.nodal)
        if ns:
            comp = np.array([s.comp for s in pb.nodal], np.int32)
            hard = np.array([int(s.hard) for s in pb.nodal], np.int32)
            box = np.ascontiguousarray(np.array([s.box for s in pb.nodal], np.int32))
            amp = np.array([s.amp for s in pb.nodal], np.float32)
            de = np.array([s.deltaevol for s in pb.nodal], np.float32)
            nu = np.array([len(s.evol) - 1 for s in pb.nodal], np.int32)
            smp = np.ascontiguousarray(np.concatenate([np.asarray(s.evol, np.float32)
                                                       for s in pb.nodal]))
            self._keep += [comp, hard, box, amp, de, nu, smp]
            lib.ref_nodal(ns, _p(comp, _i32p), _p(hard, _i32p), _p(box, _i32p), _p(amp, _f32p),
                          _p(de, _f32p), _p(nu, _i32p), _p(smp, _f32p), pb.nsteps_max)
        lib.ref_time(pb.nsteps_max)
        if len(pb.planewaves) > 1:
            raise ValueError("the reference harness drives one plane wave")
        for w in pb.planewaves:
            # the node coordinates come from the oracle's restatement of the
            # preprocessing (the reference builds them in preprocess_geom.F90)
            ln = lines(load_oracle(32), pb, 32)
            esq = np.array(w.box, np.int32)
            pp = np.ascontiguousarray(np.array(w.p, np.float32))
            ee = np.ascontiguousarray(np.array(w.e, np.float32))
            inc = np.array(w.incert, np.float32)
            ev = np.ascontiguousarray(w.evol, np.float32)
            self._keep += [ln, esq, pp, ee, inc, ev]
            lib.ref_planewave(_p(esq, _i32p), len(w.p), _p(pp, _f32p), _p(ee, _f32p),
                              _p(inc, _f32p), w.deltaevol, len(ev) - 1, _p(ev, _f32p),
                              _p(ln[0], _f32p), _p(ln[1], _f32p), _p(ln[2], _f32p))
        self.n = 0

    def tables(self):
        pb = self.pb
        nm = pb.nmedia
        f = lambda n: np.zeros(n, np.float32)
        out = [f(nm + 1) for _ in range(4)] + [f(n + 3) for n in (pb.nx, pb.ny, pb.nz) * 2]
        self.lib.ref_get_tables(*[_p(a, _f32p) for a in out])
        return out

    def step(self):
        self.lib.ref_step(self.n)
        self.n += 1

    def field(self, c):
        out = np.zeros(self.pb.alloc_shape, np.float32)
        self.lib.ref_get_field(c, _p(out, _f32p))
        return out

    def run(self, nsteps, probes=None):
        probes = self.pb.probes if probes is None else probes
        out = np.zeros((nsteps, len(probes)), np.float32)
        cache = {}
        for s in range(nsteps):
            self.step()
            cache.clear()
            for p, (c, i, j, k) in enumerate(probes):
                if c not in cache:
                    cache[c] = self.field(c)
                out[s, p] = cache[c][k + 1, j + 1, i + 1]
        return out
