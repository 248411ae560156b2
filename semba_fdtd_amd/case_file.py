"""Binary case file read by the Fortran driver (fortran/fdtd_gpu_driver.F90).

It carries the physical description of a Problem, i.e. what semba-fdtd's
``launch_simulation`` receives from the preprocessing (materials, cell sizes,
CPML parameters, poles, media matrices, nodal sources, probes).  The Fortran
side builds every coefficient table itself through the library's table
builders, in the reference's set-up order.

Layout: little-endian stream, no record markers.
    char[4] "SFDC", int32 version
    int32 nx, ny, nz, nmedia, nsteps
    float32 eps0, mu0; float64 dt
    float32 epr, sigma, mur, sigmam [nmedia+1]; int32 kind [nmedia+1]
    int32 npoles [nmedia+1]; float64 pole_c, pole_a [nmedia+1][8][2]
    float32 eps11, sig11 [nmedia+1]
    int32 mnpoles [nmedia+1]; float64 mpole_c, mpole_a [nmedia+1][8][2]
    float32 mu11, sigm11 [nmedia+1]                      (magnetic poles)
    int32 mur_faces[6]                                   (first-order Mur)
    float32 D_x [nx+4], D_y [ny+4], D_z [nz+4]        (index -2 .. n+1)
    int32 layers[6]; float32 order[6], refl[6], alphamax, alphaorden, kappamax
    int16 media index of Ex..Hz, each over its own Alloc box, x fastest
    int32 nnodal; per source: int32 comp, hard, box[6]; float32 amp, deltaevol;
        int32 numus; float32 evol[numus+1]
    int32 nprobe; int32 comp[nprobe], i[nprobe], j[nprobe], k[nprobe]
    int32 nplanewave; per wave: int32 box[6] (esqx1,esqx2,esqy1,esqy2,esqz1,
        esqz2), int32 nmodes; float32 p[nmodes][3], e[nmodes][3],
        incert[nmodes], deltaevol; int32 numus; float32 evol[numus+1]

Probe output written by the driver: int32 nsamples, nprobe; float32
samples[nsamples][nprobe].
"""
from __future__ import annotations

import numpy as np

from .problem import Problem, PMAX, to_alloc_box

MAGIC = b"SFDC"
VERSION = 3


def write_case(pb: Problem, path: str, nsteps: int) -> None:
    mi = pb.materialize()
    n3 = (pb.nx, pb.ny, pb.nz)
    nm = pb.nmedia
    i32 = lambda *v: np.asarray(v, "<i4").tobytes()
    f32 = lambda a: np.ascontiguousarray(a, "<f4").tobytes()
    f64 = lambda a: np.ascontiguousarray(a, "<f8").tobytes()
    with open(path, "wb") as f:
        f.write(MAGIC + i32(VERSION))
        f.write(i32(pb.nx, pb.ny, pb.nz, nm, nsteps))
        f.write(f32([pb.eps0, pb.mu0]) + f64([pb.dt]))
        for a in (pb.epr, pb.sigma, pb.mur, pb.sigmam):
            f.write(f32(a))
        f.write(np.ascontiguousarray(pb.kind, "<i4").tobytes())
        f.write(np.ascontiguousarray(pb.npoles, "<i4").tobytes())
        for a in (pb.pole_c, pb.pole_a):
            assert a.shape == (nm + 1, PMAX, 2)
            f.write(f64(a))
        f.write(f32(pb.eps11) + f32(pb.sig11))
        f.write(np.ascontiguousarray(pb.mnpoles, "<i4").tobytes())
        for a in (pb.mpole_c, pb.mpole_a):
            assert a.shape == (nm + 1, PMAX, 2)
            f.write(f64(a))
        f.write(f32(pb.mu11) + f32(pb.sigm11))
        f.write(i32(*pb.mur_faces))
        for a in range(3):
            f.write(f32(pb.D[a]))
        f.write(i32(*[pb.cpml_layers[a][s] for a in range(3) for s in range(2)]))
        f.write(f32([pb.cpml_order[a][s] for a in range(3) for s in range(2)]))
        f.write(f32([pb.cpml_refl[a][s] for a in range(3) for s in range(2)]))
        f.write(f32([pb.alphamax, pb.alphaorden, pb.kappamax]))
        for c in range(6):
            f.write(np.ascontiguousarray(to_alloc_box(mi[c], c, n3), "<i2").tobytes())
        f.write(i32(len(pb.nodal)))
        for s in pb.nodal:
            ev = np.ascontiguousarray(s.evol, "<f4")
            f.write(i32(s.comp, int(s.hard), *s.box))
            f.write(f32([s.amp, s.deltaevol]) + i32(len(ev) - 1) + ev.tobytes())
        P = np.array(pb.probes, "<i4").reshape(-1, 4)
        f.write(i32(len(P)))
        for col in range(4):
            f.write(np.ascontiguousarray(P[:, col]).tobytes())
        f.write(i32(len(pb.planewaves)))
        for w in pb.planewaves:
            ev = np.ascontiguousarray(w.evol, "<f4")
            f.write(i32(*w.box) + i32(len(w.p)))
            f.write(f32(np.array(w.p, np.float32).ravel()) + f32(np.array(w.e, np.float32).ravel()))
            f.write(f32(w.incert) + f32([w.deltaevol]) + i32(len(ev) - 1) + ev.tobytes())


def read_probe_file(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        ns, npr = np.frombuffer(f.read(8), "<i4")
        data = np.frombuffer(f.read(4 * int(ns) * int(npr)), "<f4")
    return data.reshape(int(ns), int(npr)).copy()
