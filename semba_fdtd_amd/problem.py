"""Problem description: everything semba-fdtd's ``launch_simulation`` receives
for the field time-stepping path, as plain numpy data.

Coordinates are total cells per axis (CPML slabs included, 0 = outer face of
the low slab).  Every field and media-index array covers the reference's Alloc
box [-1, n+1] on each axis (semba_fdtd.F90:1285-1300) and is stored x-fastest
as ``array[k+1, j+1, i+1]``.  The raw cell sizes ``D[a]`` cover [-2, n+1]
(``D[a][i+2]``), like ``sgg%dx`` (preprocess_geom.F90:5611-5666).

Media follow the reference conventions: index 0 is PEC
(preprocess_geom.F90:470-476), index 1 the background, larger indices are
bodies.  ``kind`` bits: 1 PEC, 2 PMC, 4 electric pole-residue dispersive,
8 magnetic pole-residue dispersive.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Tuple
import numpy as np

# fdetypes.F90:219-222 / semba_fdtd.F90:154-155, held in real(4)
EPS0 = float(np.float32(8.8541878176203898505365630317107502606083701665994498081024171524053950954599821142852891607182008932e-12))
MU0 = float(np.float32(1.2566370614359172953850573533118011536788677597500423283899778369231265625144835994512139301368468271e-6))
PMAX = 8
PW_MODES = 4
KIND_PEC, KIND_PMC, KIND_EDISP, KIND_MDISP = 1, 2, 4, 8
EX, EY, EZ, HX, HY, HZ = range(6)


@dataclass
class NodalSource:
    """A hard or soft line/point source on one field component
    (nodalsources.F90:245-300).  ``box`` = (xi, xe, yi, ye, zi, ze), already
    inside the component's sweep; ``evol`` = excitation samples 0..numus,
    spaced ``deltaevol`` seconds."""
    comp: int
    hard: bool
    box: Tuple[int, int, int, int, int, int]
    amp: float
    evol: np.ndarray
    deltaevol: float


@dataclass
class PlaneWave:
    """A TF/SF plane-wave box (planeonde_t, fdetypes.F90:458-465; injected by
    planewaves.F90).  ``box`` = (esqx1, esqx2, esqy1, esqy2, esqz1, esqz2) in
    total cells; per mode a propagation direction ``p``, an E amplitude ``e``
    and a start-time uncertainty ``incert`` (fraction of the box diagonal);
    ``evol`` = excitation samples 0..numus spaced ``deltaevol`` seconds."""
    box: Tuple[int, int, int, int, int, int]
    p: List[Tuple[float, float, float]]
    e: List[Tuple[float, float, float]]
    evol: np.ndarray
    deltaevol: float
    incert: List[float] = None

    def __post_init__(self):
        if self.incert is None:
            self.incert = [0.0] * len(self.p)
        assert len(self.p) == len(self.e) == len(self.incert) <= PW_MODES


@dataclass
class DftBox:
    """Frequency-domain observation of a box (observation.F90:3527-3600):
    ``kind`` 0 accumulates Ex, Ey, Ez, 1 accumulates Hx, Hy, Hz, at the
    points of ``box`` = (i1, i2, j1, j2, k1, k2) whose coordinates are
    multiples of ``stride``, for the real(4) frequencies ``freqs`` [Hz], in
    complex(8) running sums."""
    kind: int
    box: Tuple[int, int, int, int, int, int]
    stride: Tuple[int, int, int]
    freqs: np.ndarray

    def __post_init__(self):
        self.freqs = np.ascontiguousarray(self.freqs, np.float32)

    def counts(self):
        c = []
        for a in range(3):
            lo, hi, st = self.box[2 * a], self.box[2 * a + 1], self.stride[a]
            c.append(sum(1 for v in range(lo, hi + 1) if v % st == 0))
        return tuple(c)

    @property
    def shape(self):
        """(nk, nj, ni, 3, nfreq): the running sums, complex128."""
        ni, nj, nk = self.counts()
        return (nk, nj, ni, 3, len(self.freqs))


@dataclass
class Problem:
    nx: int
    ny: int
    nz: int
    epr: np.ndarray
    sigma: np.ndarray
    mur: np.ndarray
    sigmam: np.ndarray
    kind: np.ndarray
    mi: np.ndarray                      # int16 (6, nz+3, ny+3, nx+3)
    D: List[np.ndarray]                 # float32 cell sizes, [-2, n+1]
    dt: float
    npoles: np.ndarray = None
    pole_c: np.ndarray = None           # (nmedia+1, PMAX, 2) residues c_p
    pole_a: np.ndarray = None           # (nmedia+1, PMAX, 2) poles a_p
    eps11: np.ndarray = None            # absolute permittivity of dispersive media
    sig11: np.ndarray = None
    # magnetic pole-residue media (kind bit 8): poles, absolute mu, sigma_m
    mnpoles: np.ndarray = None
    mpole_c: np.ndarray = None
    mpole_a: np.ndarray = None
    mu11: np.ndarray = None
    sigm11: np.ndarray = None
    # first-order Mur faces x-, x+, y-, y+, z-, z+ (bordersmur.F90)
    mur_faces: tuple = (0, 0, 0, 0, 0, 0)
    cpml_layers: list = field(default_factory=lambda: [[0, 0], [0, 0], [0, 0]])
    cpml_order: list = field(default_factory=lambda: [[2.0, 2.0]] * 3)
    cpml_refl: list = field(default_factory=lambda: [[1e-3, 1e-3]] * 3)
    alphamax: float = 0.0
    alphaorden: float = 1.0
    kappamax: float = 1.0
    nodal: List[NodalSource] = field(default_factory=list)
    probes: List[Tuple[int, int, int, int]] = field(default_factory=list)
    planewaves: List[PlaneWave] = field(default_factory=list)
    dfts: List[DftBox] = field(default_factory=list)
    nsteps_max: int = 0
    eps0: float = EPS0
    mu0: float = MU0
    name: str = ""
    # device-side scene for grids too large to voxelise on the host: a list of
    # ("fill", m) / ("pec", faces_mask) / ("sphere", cx, cy, cz, r, m) /
    # ("box", x0, x1, y0, y1, z0, z1, m) applied in order with priorities
    # ``prio``; used when ``mi`` is None.
    scene: list = None
    prio: np.ndarray = None

    def __post_init__(self):
        nm = self.nmedia
        f32 = lambda a: np.ascontiguousarray(a, np.float32)
        self.epr, self.sigma = f32(self.epr), f32(self.sigma)
        self.mur, self.sigmam = f32(self.mur), f32(self.sigmam)
        self.kind = np.ascontiguousarray(self.kind, np.int32)
        if self.npoles is None:
            self.npoles = np.zeros(nm + 1, np.int32)
        if self.pole_c is None:
            self.pole_c = np.zeros((nm + 1, PMAX, 2))
        if self.pole_a is None:
            self.pole_a = np.zeros((nm + 1, PMAX, 2))
        if self.eps11 is None:
            self.eps11 = np.zeros(nm + 1, np.float32)
        if self.sig11 is None:
            self.sig11 = np.zeros(nm + 1, np.float32)
        self.npoles = np.ascontiguousarray(self.npoles, np.int32)
        self.eps11, self.sig11 = f32(self.eps11), f32(self.sig11)
        if self.mnpoles is None:
            self.mnpoles = np.zeros(nm + 1, np.int32)
        if self.mpole_c is None:
            self.mpole_c = np.zeros((nm + 1, PMAX, 2))
        if self.mpole_a is None:
            self.mpole_a = np.zeros((nm + 1, PMAX, 2))
        if self.mu11 is None:
            self.mu11 = np.zeros(nm + 1, np.float32)
        if self.sigm11 is None:
            self.sigm11 = np.zeros(nm + 1, np.float32)
        self.mnpoles = np.ascontiguousarray(self.mnpoles, np.int32)
        self.mu11, self.sigm11 = f32(self.mu11), f32(self.sigm11)
        self.mur_faces = tuple(int(bool(v)) for v in self.mur_faces)
        self.D = [f32(d) for d in self.D]
        if self.mi is None:
            assert self.scene is not None and self.prio is not None
        else:
            self.mi = np.ascontiguousarray(self.mi, np.int16)
            assert self.mi.shape == (6,) + self.alloc_shape, (self.mi.shape, self.alloc_shape)
        for a, n in enumerate((self.nx, self.ny, self.nz)):
            assert self.D[a].shape == (n + 4,)

    @property
    def nmedia(self) -> int:
        return len(self.epr) - 1

    @property
    def alloc_shape(self):
        return (self.nz + 3, self.ny + 3, self.nx + 3)

    @property
    def has_cpml(self) -> bool:
        return any(l for ax in self.cpml_layers for l in ax)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def materialize(self) -> np.ndarray:
        """Media index arrays of a scene-described problem, built on the host
        with the same rules the device builders apply."""
        if self.mi is not None:
            return self.mi
        n = (self.nx, self.ny, self.nz)
        mi = new_media_index(*n)
        for op in self.scene:
            kind = op[0]
            if kind == "fill":
                mi[:] = op[1]
            elif kind == "pec":
                pec_walls(mi, n, tuple(bool(op[1] >> f & 1) for f in range(6)))
            elif kind == "sphere":
                paint_voxels(mi, sphere_cells(n, op[1:4], op[4]), op[5], self.prio)
            elif kind == "box":
                x0, x1, y0, y1, z0, z1, m = op[1:]
                inside = np.zeros((self.nz, self.ny, self.nx), bool)
                inside[max(z0, 0):z1, max(y0, 0):y1, max(x0, 0):x1] = True
                paint_voxels(mi, inside, m, self.prio)
            else:
                raise ValueError(f"unknown scene op {kind}")
        return mi


# --------------------------------------------------------------------------
# sweeps (preprocess_geom.F90:5563-5591 in total coordinates)
def sweep_hi(c: int, a: int, n: Tuple[int, int, int]) -> int:
    """Last swept index of component c (0..5) along axis a."""
    node = (c != a) if c < 3 else (c - 3 == a)
    return n[a] if node else n[a] - 1


def alloc_extent(c: int, n: Tuple[int, int, int]) -> Tuple[int, int, int]:
    """(EX, EY, EZ): component c's Alloc box [-1, last swept + 1] per axis,
    the shape launch_simulation allocates it with (timestepping.F90:380-386)."""
    return tuple(sweep_hi(c, a, n) + 3 for a in range(3))


def to_alloc_box(a: np.ndarray, c: int, n) -> np.ndarray:
    """Component c's Alloc box out of a uniform (nz+3, ny+3, nx+3) array."""
    ex, ey, ez = alloc_extent(c, n)
    return np.ascontiguousarray(a[:ez, :ey, :ex])


def courant_dt(dmin, cfl=None, dt_in=None, eps0=EPS0, mu0=MU0) -> float:
    """semba_fdtd.F90:1211-1250 with real(4) arithmetic."""
    f = np.float32
    cluz = f(1.0) / np.sqrt(f(eps0) * f(mu0), dtype=np.float32)
    ix, iy, iz = (f(1.0) / f(d) for d in dmin)
    dtlay = f(1.0) / (cluz * np.sqrt(ix * ix + iy * iy + iz * iz, dtype=np.float32))
    if cfl is not None:
        return float(f(dtlay * f(cfl)))
    lim = f(dtlay * f(0.8))
    return float(lim) if (dt_in is None or dt_in > float(lim)) else float(dt_in)


def uniform_D(n: int, d: float) -> np.ndarray:
    return np.full(n + 4, d, np.float32)


def graded_D(n: int, d0: float, ratio: float, L: Tuple[int, int]) -> np.ndarray:
    """Cell sizes growing geometrically by ``ratio`` from the centre of the
    interior towards both boundaries; the CPML slabs and the padding repeat
    the edge interior cell (preprocess_geom.F90:5639-5666)."""
    D = np.empty(n + 4, np.float64)
    lo, hi = L[0], n - L[1]
    m = hi - lo
    c = (m - 1) / 2.0
    for i in range(lo, hi):
        D[i + 2] = d0 * ratio ** abs(i - c) if m > 1 else d0
    D[: lo + 2] = D[lo + 2]
    D[hi + 2:] = D[hi + 1]
    return D.astype(np.float32)


def new_media_index(nx, ny, nz, background=1) -> np.ndarray:
    return np.full((6, nz + 3, ny + 3, nx + 3), background, np.int16)


def pec_walls(mi: np.ndarray, n, faces=(True,) * 6):
    """PEC outer walls as media index 0 on tangential E and normal H
    (healer.F90:1442-1617).  faces = (xlo, xhi, ylo, yhi, zlo, zhi)."""
    nx, ny, nz = n
    def plane(c, a, v):
        sl = [slice(None)] * 3
        sl[2 - a] = v + 1
        mi[c][tuple(sl)] = 0
    for a in range(3):
        for side in range(2):
            if not faces[2 * a + side]:
                continue
            v = 0 if side == 0 else n[a]
            for c in range(3):
                if c != a:
                    plane(c, a, v)
            plane(3 + a, a, v)


def paint_voxels(mi: np.ndarray, inside: np.ndarray, medium: int, prio: np.ndarray):
    """CreateVolumeMM (healer.F90:66-234) for a set of voxels of one medium.

    ``inside`` is a bool array (nz, ny, nx) over the cells.  A voxel claims its
    12 edges when its priority is strictly larger than the edge's current
    medium (E) and its 6 faces when it is larger or equal (H)."""
    nz, ny, nx = inside.shape
    v = np.zeros((nz + 3, ny + 3, nx + 3), bool)
    v[1:nz + 1, 1:ny + 1, 1:nx + 1] = inside        # v[k+1,j+1,i+1] = voxel(i,j,k)
    def sh(dk, dj, di):
        # voxel at (i+di, j+dj, k+dk) for each array position (i, j, k)
        out = np.zeros_like(v)
        out[max(0, -dk):v.shape[0] - max(0, dk), max(0, -dj):v.shape[1] - max(0, dj),
            max(0, -di):v.shape[2] - max(0, di)] = \
            v[max(0, dk):v.shape[0] - max(0, -dk), max(0, dj):v.shape[1] - max(0, -dj),
              max(0, di):v.shape[2] - max(0, -di)]
        return out
    claim = [
        sh(0, 0, 0) | sh(0, -1, 0) | sh(-1, 0, 0) | sh(-1, -1, 0),   # Ex(i,j,k): voxels (i, j-1..j, k-1..k)
        sh(0, 0, 0) | sh(0, 0, -1) | sh(-1, 0, 0) | sh(-1, 0, -1),   # Ey: (i-1..i, j, k-1..k)
        sh(0, 0, 0) | sh(0, 0, -1) | sh(0, -1, 0) | sh(0, -1, -1),   # Ez: (i-1..i, j-1..j, k)
        sh(0, 0, 0) | sh(0, 0, -1),                                   # Hx: (i-1..i, j, k)
        sh(0, 0, 0) | sh(0, -1, 0),                                   # Hy: (i, j-1..j, k)
        sh(0, 0, 0) | sh(-1, 0, 0),                                   # Hz: (i, j, k-1..k)
    ]
    p_new = prio[medium]
    for c in range(6):
        cur = prio[mi[c].astype(np.int64)]
        win = (p_new > cur) if c < 3 else (p_new >= cur)
        mi[c][claim[c] & win] = medium


def sphere_cells(n, centre, radius) -> np.ndarray:
    """Voxels whose centre lies inside the sphere (cells indexed (k, j, i)),
    evaluated in real(4) like the device builder."""
    nx, ny, nz = n
    f = np.float32
    k, j, i = np.meshgrid(np.arange(nz, dtype=f) + f(0.5), np.arange(ny, dtype=f) + f(0.5),
                          np.arange(nx, dtype=f) + f(0.5), indexing="ij")
    cx, cy, cz = (f(c) for c in centre)
    dx, dy, dz = i - cx, j - cy, k - cz
    r = f(radius)
    return dx * dx + dy * dy + dz * dz <= r * r


def gaussian_table(dt: float, nsamples: int, spread: float, delay: float) -> np.ndarray:
    """Excitation samples g(n) = exp(-((n - delay)/spread)^2) at spacing dt."""
    n = np.arange(nsamples + 1, dtype=np.float64)
    return np.exp(-((n - delay) / spread) ** 2).astype(np.float32)


# --------------------------------------------------------------------------
def cavity_problem(n: int = 64, nsteps: int = 2000, seed: int = 1001, hard: bool = True,
                   dx: float = 0.01, nprobes: int = 8) -> Problem:
    """Plain method, CPU-runnable: n^3 vacuum cavity, PEC walls, Hertzian Ez
    dipole at a seeded interior edge driven by a Gaussian (spread 30 steps,
    delay 120 steps), dt = 0.9 x Courant, ``nprobes`` seeded Ex/Ey/Ez probes."""
    rng = np.random.default_rng(seed)
    mi = new_media_index(n, n, n)
    pec_walls(mi, (n, n, n))
    dt = courant_dt((dx, dx, dx), cfl=0.9)
    s = rng.integers(n // 4, 3 * n // 4, size=3)
    evol = gaussian_table(dt, nsteps + 4, 30.0, 120.0)
    src = NodalSource(EZ, hard, (int(s[0]), int(s[0]), int(s[1]), int(s[1]), int(s[2]), int(s[2])),
                      1.0, evol, float(np.float32(dt)))
    probes = []
    for p in range(nprobes):
        c = int(rng.integers(0, 3))
        pos = rng.integers(2, n - 2, size=3)
        probes.append((c, int(pos[0]), int(pos[1]), int(pos[2])))
    return Problem(n, n, n, epr=[1.0, 1.0], sigma=[0.0, 0.0], mur=[1.0, 1.0],
                   sigmam=[0.0, 0.0], kind=[KIND_PEC, 0], mi=mi,
                   D=[uniform_D(n, dx)] * 3, dt=dt, nodal=[src], probes=probes,
                   nsteps_max=nsteps, name=f"cavity{n}")


def lorentz_pair(f0: float, delta: float, deps: float, eps0: float = EPS0):
    """A complex-conjugate pole pair with resonance f0 [Hz], damping delta
    [1/s] and static increment deps (relative).  Both poles are listed
    explicitly, as the reference requires (electricdispersive.F90:26-31).
    Returns (c, a) arrays of shape (2, 2) = [pole][re, im]."""
    a = complex(-delta, 2 * np.pi * f0)
    c = -deps * eps0 * a / 2.0            # -2 Re(c/a) = deps*eps0 at s = 0
    cs = np.array([[c.real, c.imag], [c.real, -c.imag]])
    as_ = np.array([[a.real, a.imag], [a.real, -a.imag]])
    return cs, as_


def heterogeneous_problem(n: int = 1024, L: int = 12, seed: int = 1003, nsteps: int = 2000,
                          dx: float = 0.01, nspheres: int = 8, disp_fraction: float = 0.2,
                          nprobes: int = 16, host_media: bool = False) -> Problem:
    """Large heterogeneous single-GPU grid: n^3 total cells with an L-cell
    CPML on every face, ``nspheres`` seeded voxelised spheres (radius
    20-120 cells at n = 1024, scaled with n) each of one of 8 lossy
    dielectrics (eps_r U[1,12], sigma U[0,0.5] S/m), and a 2-pole
    complex-conjugate pole-residue medium (resonance U[1,10] GHz) filling
    ``disp_fraction`` of the volume as a slab inside the CPML.  Excited by a
    TF/SF Gaussian plane wave (+z incidence, Ex polarisation) on a box one
    cell inside the CPML; the bodies stay clear of its faces.  The media
    matrices are described as a scene for the device builders
    (``host_media`` materialises them on the host instead)."""
    rng = np.random.default_rng(seed)
    s = n / 1024.0
    nd = 8
    epr = [1.0, 1.0] + list(rng.uniform(1.0, 12.0, nd))
    sig = [0.0, 0.0] + list(rng.uniform(0.0, 0.5, nd))
    kind = [KIND_PEC, 0] + [0] * nd
    mdisp = len(epr)
    epr.append(2.0); sig.append(0.0); kind.append(KIND_EDISP)
    nm = len(epr) - 1
    f0 = float(rng.uniform(1e9, 10e9))
    c, a = lorentz_pair(f0, 0.1 * 2 * np.pi * f0, 2.0)
    pc = np.zeros((nm + 1, PMAX, 2)); pa = np.zeros((nm + 1, PMAX, 2))
    pc[mdisp, :2] = c; pa[mdisp, :2] = a
    npoles = np.zeros(nm + 1, np.int32); npoles[mdisp] = 2
    eps11 = np.zeros(nm + 1, np.float32); eps11[mdisp] = np.float32(2.0 * EPS0)
    sig11 = np.zeros(nm + 1, np.float32)
    prio = np.array([1000, 0] + [20 + i for i in range(nd)] + [10], np.int64)
    lo, hi = L, n - L                 # interior cells [lo, hi)
    blo, bhi = lo + 1, hi - 1         # TF/SF box (cells)
    inner = hi - lo
    thick = max(1, int(round(disp_fraction * n ** 3 / (inner - 4) ** 2))) if disp_fraction > 0 else 0
    zc = (lo + hi) // 2
    scene = [("fill", 1)]
    if thick:
        z0 = max(blo + 2, zc - thick // 2)
        scene.append(("box", blo + 2, bhi - 2, blo + 2, bhi - 2, z0, min(bhi - 2, z0 + thick),
                      mdisp))
    for b in range(nspheres):
        r = max(2.0, float(rng.uniform(20.0, 120.0)) * s)
        r = min(r, 0.4 * inner)
        ctr = [float(rng.uniform(blo + r + 3, bhi - r - 3)) for _ in range(3)]
        scene.append(("sphere", ctr[0], ctr[1], ctr[2], r, 2 + b % nd))
    D = [uniform_D(n, dx)] * 3
    dt = courant_dt((dx, dx, dx), cfl=0.9)
    evol = gaussian_table(dt, nsteps + 4, 30.0, 120.0)
    pw = PlaneWave((blo, bhi, blo, bhi, blo, bhi), [(0.0, 0.0, 1.0)], [(1.0, 0.0, 0.0)], evol,
                   float(np.float32(dt)))
    probes = []
    for p in range(nprobes):
        cc = p % 6
        probes.append((cc,) + tuple(int(rng.integers(blo + 1, bhi - 1)) for _ in range(3)))
    pb = Problem(n, n, n, epr=epr, sigma=sig, mur=[1.0] * (nm + 1), sigmam=[0.0] * (nm + 1),
                 kind=kind, mi=None, D=D, dt=dt, npoles=npoles, pole_c=pc, pole_a=pa,
                 eps11=eps11, sig11=sig11, cpml_layers=[[L, L]] * 3,
                 cpml_order=[[3.0, 3.0]] * 3, cpml_refl=[[1e-6, 1e-6]] * 3, planewaves=[pw],
                 probes=probes, nsteps_max=nsteps, name=f"heterogeneous{n}", scene=scene,
                 prio=prio)
    if host_media:
        pb.mi = pb.materialize()
    return pb


def scatterer_problem(n=(24, 22, 20), L=((4, 5), (3, 4), (5, 3)), seed: int = 7,
                      nsteps: int = 300, dx: float = 0.01, graded_z: float = 1.0,
                      order: float = 3.0, refl: float = 1e-6, disp: bool = True,
                      soft: bool = True, alphamax: float = 0.0, kappamax: float = 1.0,
                      nprobes: int = 12, mdisp: bool = False, mur=None) -> Problem:
    """Small open problem exercising every piece of the path: CPML of uneven
    thickness on all faces, optional geometric z grading, a lossy dielectric
    sphere, a 2-pole Lorentz blob, a soft Ex and a hard Hy source and probes
    on all six components.  ``mdisp`` adds a magnetic 2-pole Lorentz block
    and makes the Lorentz blob magnetic-dispersive too (one real pole);
    ``mur`` = 6 flags (x-, x+, y-, y+, z-, z+) of first-order Mur faces."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    mi = new_media_index(nx, ny, nz)
    epr = [1.0, 1.0, 4.0]
    sig = [0.0, 0.0, 0.01]
    kind = [KIND_PEC, 0, 0]
    prio = np.array([100, 0, 10, 20], np.int64)
    inner = [(L[a][0], (nx, ny, nz)[a] - L[a][1]) for a in range(3)]
    ctr = [(lo + hi) / 2.0 for lo, hi in inner]
    rad = 0.3 * min(hi - lo for lo, hi in inner)
    paint_voxels(mi, sphere_cells(n, ctr, rad), 2, prio)
    npoles = [0, 0, 0]
    pc = np.zeros((6, PMAX, 2)); pa = np.zeros((6, PMAX, 2))
    eps11 = np.zeros(6, np.float32); sig11 = np.zeros(6, np.float32)
    if disp:
        epr.append(2.0); sig.append(0.005); kind.append(KIND_EDISP); npoles.append(2)
        c, a = lorentz_pair(5e9, 2e9, 1.5)
        pc[3, :2] = c; pa[3, :2] = a
        eps11[3] = np.float32(2.0 * EPS0); sig11[3] = 0.005
        box = np.zeros((nz, ny, nx), bool)
        x0, y0, z0 = (inner[0][0] + 1, inner[1][0] + 1, inner[2][0] + 1)
        box[z0:z0 + 4, y0:y0 + 5, x0:x0 + 6] = True
        paint_voxels(mi, box, 3, prio)
    mnp = [0] * 6
    mpc = np.zeros((6, PMAX, 2)); mpa = np.zeros((6, PMAX, 2))
    mu11 = np.zeros(6, np.float32); sigm11 = np.zeros(6, np.float32)
    mur_rel = [1.0] * 6
    if mdisp:
        if disp:  # the electric Lorentz blob also carries one real magnetic pole
            kind[3] |= KIND_MDISP
            mnp[3] = 1
            mpc[3, 0] = (0.4 * MU0 * 2e9, 0.0); mpa[3, 0] = (-2e9, 0.0)
            mu11[3] = np.float32(MU0); sigm11[3] = 0.0
        m = len(epr)
        epr.append(1.0); sig.append(0.0); kind.append(KIND_MDISP)
        prio = np.concatenate([prio[:m], [30]]).astype(np.int64)
        mnp[m] = 2
        c, a = lorentz_pair(3e9, 1e9, 0.8, eps0=MU0)
        mpc[m, :2] = c; mpa[m, :2] = a
        mu11[m] = np.float32(1.5 * MU0); sigm11[m] = 40.0
        mur_rel[m] = 1.5
        box = np.zeros((nz, ny, nx), bool)
        x1, y1, z1 = inner[0][1] - 1, inner[1][1] - 1, inner[2][1] - 1
        box[z1 - 4:z1, y1 - 5:y1, x1 - 6:x1] = True
        paint_voxels(mi, box, m, prio)
    nm = len(epr) - 1
    D = [uniform_D(nx, dx), uniform_D(ny, dx),
         graded_D(nz, dx, graded_z, L[2]) if graded_z != 1.0 else uniform_D(nz, dx)]
    dmin = [float(d.min()) for d in D]
    dt = courant_dt(dmin, cfl=0.9)
    evol = gaussian_table(dt, nsteps + 4, 12.0, 40.0)
    s = [int(rng.integers(inner[a][0] + 2, inner[a][1] - 2)) for a in range(3)]
    srcs = [NodalSource(EX, not soft, (s[0], s[0], s[1], s[1], s[2], s[2]), 3.0, evol,
                        float(np.float32(dt))),
            NodalSource(HY, True, (s[0] + 1, s[0] + 1, s[1], s[1], s[2] - 1, s[2] - 1), 0.01,
                        evol, float(np.float32(dt)))]
    probes = []
    for p in range(nprobes):
        c = p % 6
        probes.append((c,) + tuple(int(rng.integers(0, sweep_hi(c, a, n) + 1)) for a in range(3)))
    npoles += [0] * (nm + 1 - len(npoles))
    return Problem(nx, ny, nz, epr=epr, sigma=sig, mur=mur_rel[:nm + 1], sigmam=[0.0] * (nm + 1),
                   kind=kind, mi=mi, D=D, dt=dt, npoles=npoles, pole_c=pc[:nm + 1],
                   pole_a=pa[:nm + 1], eps11=eps11[:nm + 1], sig11=sig11[:nm + 1],
                   mnpoles=mnp[:nm + 1], mpole_c=mpc[:nm + 1], mpole_a=mpa[:nm + 1],
                   mu11=mu11[:nm + 1], sigm11=sigm11[:nm + 1],
                   mur_faces=tuple(mur) if mur else (0,) * 6,
                   cpml_layers=[list(l) for l in L], cpml_order=[[order, order]] * 3,
                   cpml_refl=[[refl, refl]] * 3, alphamax=alphamax, kappamax=kappamax,
                   nodal=srcs, probes=probes, nsteps_max=nsteps, name="scatterer")


def open_scatterer_problem(n: int = 256, L: int = 10, radius: float = 40.0, grading: float = 1.03,
                           nsteps: int = 4000, dx: float = 0.01, margin: int = 3, eps_r: float = 4.0,
                           sigma: float = 0.01, nprobes: int = 16, seed: int = 1002,
                           spread: float = 30.0, delay: float = 120.0,
                           direction=(0.0, 0.0, 1.0), polar=(1.0, 0.0, 0.0)) -> Problem:
    """Open scatterer: n^3 total cells with an L-cell CPML on every face
    (grading m = 3, R0 = 1e-6, kappa_max = 1, alpha = 0), a centred dielectric
    sphere (radius in cells, eps_r, sigma), the z mesh graded geometrically by
    ``grading`` from the centre outwards, and a TF/SF Gaussian plane wave
    (+z incidence, Ex polarisation by default) on a box ``margin`` cells inside
    the CPML.  ``nprobes`` seeded probes on all six components."""
    rng = np.random.default_rng(seed)
    mi = new_media_index(n, n, n)
    prio = np.array([100, 0, 10], np.int64)
    ctr = [n / 2.0] * 3
    paint_voxels(mi, sphere_cells((n, n, n), ctr, radius), 2, prio)
    D = [uniform_D(n, dx), uniform_D(n, dx),
         graded_D(n, dx, grading, (L, L)) if grading != 1.0 else uniform_D(n, dx)]
    dt = courant_dt([float(d.min()) for d in D], cfl=0.9)
    lo, hi = L + margin, n - L - margin
    evol = gaussian_table(dt, nsteps + 4, spread, delay)
    pw = PlaneWave((lo, hi, lo, hi, lo, hi), [tuple(direction)], [tuple(polar)], evol,
                   float(np.float32(dt)))
    probes = []
    for p in range(nprobes):
        c = p % 6
        probes.append((c,) + tuple(int(rng.integers(L + 1, n - L - 1)) for _ in range(3)))
    return Problem(n, n, n, epr=[1.0, 1.0, eps_r], sigma=[0.0, 0.0, sigma], mur=[1.0] * 3,
                   sigmam=[0.0] * 3, kind=[KIND_PEC, 0, 0], mi=mi, D=D, dt=dt,
                   cpml_layers=[[L, L]] * 3, cpml_order=[[3.0, 3.0]] * 3,
                   cpml_refl=[[1e-6, 1e-6]] * 3, planewaves=[pw], probes=probes,
                   nsteps_max=nsteps, name=f"open_scatterer{n}")


def ellipsoid_cells(n, centre, axes) -> np.ndarray:
    """Cells whose centre lies inside the axis-aligned ellipsoid, in real(4)."""
    nx, ny, nz = n
    f = np.float32
    k, j, i = np.meshgrid(np.arange(nz, dtype=f) + f(0.5), np.arange(ny, dtype=f) + f(0.5),
                          np.arange(nx, dtype=f) + f(0.5), indexing="ij")
    q = [((v - f(c)) / f(a)) for v, c, a in zip((i, j, k), centre, axes)]
    return q[0] * q[0] + q[1] * q[1] + q[2] * q[2] <= f(1.0)


def blob_problem(n, nsteps: int, rng, L: int = 8, nblobs=(1, 6), eps_range=(1.0, 10.0),
                 dispersive: bool = False, dx: float = 0.01, nprobes: int = 4,
                 nfreq: int = 0, name: str = "blobs") -> Problem:
    """An independent scattering solve: n = (nx, ny, nz) total cells with an
    L-cell CPML, a few seeded dielectric ellipsoids (eps_r U[eps_range]), an
    optional 1-pole (Debye-like) dispersive box, a TF/SF Gaussian plane wave
    (+z, Ex) two cells inside the CPML, point probes and, with nfreq > 0, a
    frequency-domain box over the probe region at nfreq frequencies."""
    nx, ny, nz = n
    mi = new_media_index(nx, ny, nz)
    nb = int(rng.integers(nblobs[0], nblobs[1] + 1))
    epr = [1.0, 1.0] + [float(rng.uniform(*eps_range)) for _ in range(nb)]
    sig = [0.0] * len(epr)
    kind = [KIND_PEC, 0] + [0] * nb
    prio = [100, 0] + [10 + b for b in range(nb)]
    lo = [L + 2] * 3
    hi = [v - L - 2 for v in n]
    for b in range(nb):
        axes = [float(rng.uniform(2.0, max(2.5, 0.25 * (h - l)))) for l, h in zip(lo, hi)]
        ctr = [float(rng.uniform(l + a + 1, max(l + a + 1.5, h - a - 1)))
               for l, h, a in zip(lo, hi, axes)]
        paint_voxels(mi, ellipsoid_cells(n, ctr, axes), 2 + b, np.array(prio + [0], np.int64))
    npoles = [0] * len(epr)
    pc = np.zeros((len(epr) + 1, PMAX, 2)); pa = np.zeros((len(epr) + 1, PMAX, 2))
    eps11 = np.zeros(len(epr) + 1, np.float32); sig11 = np.zeros(len(epr) + 1, np.float32)
    if dispersive:
        md = len(epr)
        epr.append(2.0); sig.append(0.0); kind.append(KIND_EDISP); npoles.append(1)
        prio.append(5)
        a = -2.0 * np.pi * float(rng.uniform(1e9, 10e9))     # real pole
        pa[md, 0] = (a, 0.0)
        pc[md, 0] = (-1.5 * EPS0 * a, 0.0)                     # static increment 1.5
        eps11[md] = np.float32(2.0 * EPS0)
        box = np.zeros((nz, ny, nx), bool)
        box[lo[2] + 1:hi[2] - 1, lo[1] + 1:(lo[1] + hi[1]) // 2, lo[0] + 1:hi[0] - 1] = True
        paint_voxels(mi, box, md, np.array(prio, np.int64))
    nm = len(epr) - 1
    D = [uniform_D(v, dx) for v in n]
    dt = courant_dt((dx, dx, dx), cfl=0.9)
    evol = gaussian_table(dt, nsteps + 4, 20.0, 80.0)
    pw = PlaneWave((L + 2, nx - L - 2, L + 2, ny - L - 2, L + 2, nz - L - 2), [(0.0, 0.0, 1.0)],
                   [(1.0, 0.0, 0.0)], evol, float(np.float32(dt)))
    probes = []
    for p in range(nprobes):
        probes.append((p % 6,) + tuple(int(rng.integers(lo[a], hi[a])) for a in range(3)))
    dfts = []
    if nfreq > 0:
        c = [v // 2 for v in n]
        freqs = np.linspace(0.5e9, 8e9, nfreq).astype(np.float32)
        dfts.append(DftBox(0, (c[0] - 1, c[0] + 1, c[1] - 1, c[1] + 1, c[2], c[2]), (1, 1, 1),
                           freqs))
    return Problem(nx, ny, nz, epr=epr, sigma=sig, mur=[1.0] * (nm + 1), sigmam=[0.0] * (nm + 1),
                   kind=kind, mi=mi, D=D, dt=dt, npoles=np.array(npoles[:nm + 1]),
                   pole_c=pc[:nm + 1], pole_a=pa[:nm + 1], eps11=eps11[:nm + 1],
                   sig11=sig11[:nm + 1], cpml_layers=[[L, L]] * 3, cpml_order=[[3.0, 3.0]] * 3,
                   cpml_refl=[[1e-6, 1e-6]] * 3, planewaves=[pw], probes=probes, dfts=dfts,
                   nsteps_max=nsteps, name=name)


def batch_instances(count: int = 256, nsteps: int = 1500, seed: int = 1004, lo: int = 32,
                    hi: int = 128):
    """Batched uneven instances: edge lengths log-uniform in [lo, hi] per
    axis, 8-cell CPML, 1-6 ellipsoidal blobs (eps_r U[1,10]), a plane wave."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = tuple(int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3))
        out.append(blob_problem(n, nsteps, rng, name=f"instance{q}"))
    return out


def serving_requests(count: int = 48, rate: float = 8.0, seed: int = 1005, nmin: int = 48,
                     nmax: int = 192, steps=(500, 4000), nfreq: int = 64,
                     disp_share: float = 0.3):
    """Serving stream: Poisson arrivals at ``rate`` per second of cubic solves
    U{nmin..nmax}^3, U[steps] steps, 8-cell CPML, random blobs, a 1-pole
    dispersive box in ``disp_share`` of the requests and a 64-frequency
    FP64 DFT over the centre.  Returns [(arrival_s, Problem, nsteps)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(rng.integers(nmin, nmax + 1))
        ns = int(rng.integers(steps[0], steps[1] + 1))
        pb = blob_problem((n, n, n), ns, rng, dispersive=bool(rng.random() < disp_share),
                          nfreq=nfreq, name=f"request{q}")
        out.append((t, pb, ns))
    return out
