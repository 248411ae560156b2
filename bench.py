"""Throughput of the gfx950 time-stepping path on one MI355X.

Workload (default): the large heterogeneous single-GPU grid -- 1024^3 total
cells with a 12-cell CPML on every face, 8 seeded lossy dielectric spheres and
a 2-pole pole-residue medium in 20% of the volume -- built in HBM by the
device scene builders.  A step is one full Yee iteration (E half + H half,
CPML and pole state fused in, probes sampled), replayed from a captured HIP
graph.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--n N] [--mode fast|exact]
                    [--kernel-reps R] [--no-cpu] [--cpu-n N] [--cpu-steps K]
                    [--no-batch] [--batch-instances N] [--batch-steps K]
                    [--no-serve] [--serve-requests N] [--serve-rate R]

Besides the 1024^3 line it measures the batched-instances workload (256
uneven grids, one persistent launch per half step, against the plain
one-at-a-time path) and a serving stream (continuous batching against the
plain path: p50/p99 latency and aggregate Mcell-updates/s), reported in the
"batch" and "serve" objects of the same JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
SERVE_MAX_CELLS = 24_000_000  # cells admitted at once by the batched server
METRIC = "Mcell-updates/s + %HBM roofline; batched-serve p50/p99 latency vs plain"


def psi_counts(pb):
    """psi values updated per half step (E side, H side): the CPML slab
    arrays of both passes."""
    nx, ny, nz = pb.nx, pb.ny, pb.nz
    Lx, Ly, Lz = (sum(pb.cpml_layers[a]) for a in range(3))
    e = (nx * Ly * (nz + 1) + nx * (ny + 1) * Lz + Lx * ny * (nz + 1) + (nx + 1) * ny * Lz
         + Lx * (ny + 1) * nz + (nx + 1) * Ly * nz)
    h = ((nx + 1) * Ly * nz + (nx + 1) * ny * Lz + Lx * (ny + 1) * nz + nx * (ny + 1) * Lz
         + Lx * ny * (nz + 1) + nx * Ly * (nz + 1))
    return e, h


def algorithmic_bytes(pb, ndisp_edges, mode, stored_poles):
    """Compulsory HBM bytes per launch of the three step kernels.

    E face pass (Yee E + CPML E) and H pass (Yee H + CPML H), per cell: read
    the 3 updated components (12 B), the 3 curl components (12 B) and 3 int16
    media indices (6 B), write 3 components (12 B) = 42 B; CPML adds 8 B per
    psi value (read + write).  Pole pass, per dispersive edge: E read + write
    (8 B), Eprev read + write (8 B) and 16 B per stored complex64 pole state
    (FAST, conjugate pairs folded) or 32 B per complex128 state (EXACT)."""
    cells = pb.cells
    pe, ph = psi_counts(pb)
    per_state = 16 if mode == "fast" else 32
    return {"e_face": 42 * cells + 8 * pe,
            "ade": ndisp_edges * (16 + per_state * stored_poles),
            "h": 42 * cells + 8 * ph}


def step_algorithmic_bytes(pb, ndisp_edges, mode, stored_poles):
    """Compulsory HBM bytes of one whole step with half-step passes that touch
    each array once: 84 B per cell (two passes of 42 B), 8 B per psi value,
    and per dispersive edge 8 B of Eprev plus the pole state (its E access is
    already in the 84 B)."""
    pe, ph = psi_counts(pb)
    per_state = 16 if mode == "fast" else 32
    return 84 * pb.cells + 8 * (pe + ph) + ndisp_edges * (8 + per_state * stored_poles)


def media_bytes(pb):
    """Bytes per media index the one-launch step reads: the library keeps
    8-bit copies of the media arrays for it when there are at most 256 media."""
    return 1 if pb.nmedia + 1 <= 256 else 2


def onelaunch_cell_bytes(pb):
    """Compulsory bytes per cell of the one-launch step: each of the 6 fields
    read once and written once, each of the 6 media indices read once."""
    return 48 + 6 * media_bytes(pb)


def onelaunch_step_bytes(pb, ndisp_edges, mode, stored_poles):
    """Compulsory HBM bytes of one whole step when E and H of a plane are
    updated in the same launch: onelaunch_cell_bytes per cell, 8 B per psi
    value and per dispersive edge 8 B of Eprev plus the pole state."""
    pe, ph = psi_counts(pb)
    per_state = 16 if mode == "fast" else 32
    return (onelaunch_cell_bytes(pb) * pb.cells + 8 * (pe + ph)
            + ndisp_edges * (8 + per_state * stored_poles))


CLASS_NAMES = {0: "plain", 1: "cpml", 2: "poles", 3: "cpml+poles", 4: "sources",
               5: "cpml+sources", 6: "poles+sources", 7: "cpml+poles+sources"}

KERNELS = {
    "e_face": "k_epass (Yee E + CPML E, fused face pass)",
    "ade": "k_ade_box (pole-residue ADE over the dispersive edges)",
    "h": "k_hpass (Yee H + CPML H, fused face pass)",
    "step_plain": "k_step<R, false, false, false, media type> (one launch per step, interior "
                  "tiles: Yee E and H of each plane)",
}


def cpu_baseline(n_sample, steps):
    """The oracle (C restatement, OpenMP over k/j like the reference loops)
    on the host cores, on the same heterogeneous workload at n_sample^3."""
    import oracle
    from semba_fdtd_amd import problem as P
    pb = P.heterogeneous_problem(n=n_sample, L=max(4, 12 * n_sample // 1024), nsteps=steps,
                                 host_media=True)
    o = oracle.Oracle(pb)
    o.run(2)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(pb.cells * steps / dt / 1e6, 2), "unit": "Mcell-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"oracle (fdtd_oracle.c, -O2 OpenMP {threads} threads) on the same "
                      f"heterogeneous workload at {n_sample}^3 (CPML {pb.cpml_layers[0][0]}, "
                      f"spheres, 20% 2-pole medium, complex128 poles), {steps} steps, "
                      f"{dt:.1f} s"}


def batch_leg(count, steps, plain_sample, mode):
    """Batched uneven instances (log-uniform 32-128 cells per axis, 8-cell
    CPML, 1-6 blobs, plane wave): all instances advanced together, against a
    sample of them solved one at a time on their own graphs."""
    import torch
    import semba_fdtd_amd as sf
    from semba_fdtd_amd import problem as P
    from semba_fdtd_amd.batch import Batch
    insts = P.batch_instances(count=count, nsteps=steps)
    cells = sum(p.cells for p in insts)
    print(f"[bench] batch: {count} instances, {cells / 1e6:.1f} M cells, {steps} steps",
          file=sys.stderr, flush=True)
    sol = [sf.Solver(p, mode=mode, probe_buffer=steps) for p in insts]
    b = Batch()
    for s in sol:
        b.admit(s)
    b.step(2)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps - 2)
    b.synchronize()
    tb = time.perf_counter() - t0
    batched = cells * (steps - 2) / tb / 1e6
    finite = all(bool(np.isfinite(s.probes()).all()) for s in sol)
    b.close()
    for s in sol:
        s.close()
    del sol
    # plain path on a sample: each instance alone, one after the other
    sample = insts[:plain_sample]
    scells = sum(p.cells for p in sample)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in sample:
        s = sf.Solver(p, mode=mode, probe_buffer=steps)
        s.run(steps)
        s.synchronize()
        s.close()
    tp = time.perf_counter() - t0
    plain = scells * steps / tp / 1e6
    return {"instances": count, "cells": cells, "steps": steps, "mode": mode,
            "batched_mcell_updates_per_s": round(batched, 1),
            "plain_mcell_updates_per_s": round(plain, 1),
            "plain_sample_instances": len(sample),
            "speedup": round(batched / plain, 2), "probes_finite": finite,
            "note": "batched: one persistent face-pass launch per half step for all instances, "
                    "graph-replayed; plain: each instance alone on its own graph (set-up "
                    "included), throughput measured on the first instances"}


def serve_leg(count, rate, mode):
    """Serving stream (Poisson arrivals, cubic 48-192, 500-4000 steps, 30%
    with a 1-pole region, 64-frequency FP64 DFT): continuous batching against
    the plain one-solve-at-a-time path on the same stream."""
    from semba_fdtd_amd import problem as P
    from semba_fdtd_amd import serve
    reqs = P.serving_requests(count=count, rate=rate)
    print(f"[bench] serve: {count} requests at {rate}/s", file=sys.stderr, flush=True)
    plain = serve.serve_plain(reqs, mode=mode)
    # at most SERVE_MAX_CELLS cells admitted at once: enough to fill the
    # device, few enough that early requests finish early
    batched = serve.serve_batched(reqs, mode=mode, max_cells=SERVE_MAX_CELLS)
    same = True
    pm = {r.name: r for r in plain}
    for r in batched:
        same = same and np.array_equal(r.probes.view(np.uint32), pm[r.name].probes.view(np.uint32))
    return {"requests": count, "arrival_rate_per_s": rate, "mode": mode,
            "batched_max_cells": SERVE_MAX_CELLS,
            "batched": serve.summarize(batched), "plain": serve.summarize(plain),
            "results_identical": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--mode", default="fast", choices=["fast", "exact"])
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--disp-fraction", type=float, default=0.2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 plain 1 GiB device copies (the known-size kernel that "
                         "scales the FETCH_SIZE/WRITE_SIZE counters) and exit")
    ap.add_argument("--tiling", default="1,16,1,16,16",
                    help="work decomposition: rows,zchunk,xcells of the two passes and "
                         "rows,zchunk of the one-launch step (step rows 0: two passes per "
                         "step); the default is the fastest measured on this workload")
    ap.add_argument("--tune", action="store_true",
                    help="time the candidate decompositions on the grid and keep the fastest "
                         "instead of --tiling")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=1500)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only run semba_fdtd_amd/hbm_probe (achievable HBM bandwidth per access "
                         "shape) and exit")
    args = ap.parse_args()
    if args.hbm_probe:
        import subprocess
        exe = os.path.join(ROOT, "semba_fdtd_amd", "hbm_probe")
        sys.exit(subprocess.run([exe]).returncode)
    if args.copy_calibration:
        import torch
        x = torch.ones(2**28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read_per_launch": x.numel() * 4,
                          "bytes_written_per_launch": y.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: the "
              f"whole grid is resident on one GPU and nothing here spans several",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import semba_fdtd_amd as sf
    from semba_fdtd_amd import problem as P

    total = args.warmup + args.steps + 2 * args.kernel_reps + 4
    pb = P.heterogeneous_problem(n=args.n, nsteps=total, disp_fraction=args.disp_fraction)
    print(f"[bench] building {pb.name} ({pb.cells / 1e9:.3f} G cells) mode={args.mode}",
          file=sys.stderr, flush=True)
    torch.cuda.init()
    stream = torch.cuda.Stream()
    t0 = time.perf_counter()
    s = sf.Solver(pb, mode=args.mode, stream=stream.cuda_stream, probe_buffer=total)
    if args.tune:
        # the grid is at rest: the source-free passes leave it at zero, so the
        # sweep needs no checkpoint
        from semba_fdtd_amd import tuning
        tuned = tuning.autotune(s, reps=3, keep_state=False)
        tiling = {"passes": list(tuned["tiling"]), "step": list(tuned["step"]), "tuned": True}
        if tuned["step"][0]:
            tiling["class_rows"] = s.get_step_class_rows()
    else:
        t = [int(v) for v in args.tiling.split(",")]
        s.set_tiling(t[0], t[1], t[2])
        s.set_step_mode(t[3], t[4])
        tiling = {"passes": t[:3], "step": t[3:5], "tuned": False}
        if t[3]:
            tiling["class_rows"] = s.get_step_class_rows()
    print(f"[bench] tiling {tiling}", file=sys.stderr, flush=True)
    s.run(args.warmup)  # builds pole state, captures the graph
    s.synchronize()
    setup_s = time.perf_counter() - t0
    ndisp = sum(int(s.lib.fdtd_dispersive_edges(s.h, c)) for c in range(3))
    print(f"[bench] setup {setup_s:.1f} s, {s.device_bytes / 2**30:.1f} GiB in HBM, "
          f"{ndisp} dispersive edges", file=sys.stderr, flush=True)

    # timed region: K whole steps
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        ev0.record(stream)
        s.run(args.steps)
        ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    ms_per_step = ms / args.steps
    value = pb.cells * args.steps / (ms / 1e3) / 1e6

    # per-kernel durations on the stream the kernels run on.  fdtd_advance_h
    # launches k_hpass alone (this workload has no H sources),
    # fdtd_advance_e_dispersive launches k_ade_box alone, and fdtd_advance_e
    # launches k_epass then k_ade_box (+ the nodal E kernel).
    def timed(fn, reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps
    lib, h = s.lib, s.h
    n0 = s.n
    reps = args.kernel_reps
    ms_k = {"h": timed(lambda: lib.fdtd_advance_h(h, n0), reps)}
    ms_e = timed(lambda: lib.fdtd_advance_e(h, n0), reps)
    ms_k["ade"] = timed(lambda: lib.fdtd_advance_e_dispersive(h), reps) if ndisp else 0.0
    ms_k["e_face"] = ms_e - ms_k["ade"]
    stored = 1 if args.mode == "fast" else 2
    nbytes = algorithmic_bytes(pb, ndisp, args.mode, stored)
    onelaunch = bool(s.get_step_mode()[0])
    classes = None
    if onelaunch:
        # one launch per step: each tile class's kernel timed alone; the
        # interior (plain) class moves exactly onelaunch_cell_bytes per owned
        # cell
        ms_k = {}
        classes = []
        cb = onelaunch_cell_bytes(pb)
        for cls, cells, ms_c in s.step_classes(reps):
            key = "step_" + CLASS_NAMES[cls].replace("+", "_")
            ms_k[key] = ms_c
            nbytes[key] = int(cb * cells)
            classes.append({"class": CLASS_NAMES[cls], "cells": int(cells), "ms": round(ms_c, 4)})
        dom = max(ms_k, key=ms_k.get)
        KERNELS["step_plain"] += f", {cb} B per owned cell"
        if dom != "step_plain":
            KERNELS[dom] = (f"k_step tile class {dom[5:]} ({cb} B per owned cell counted: a lower "
                            f"bound of its compulsory bytes)")
        step_bytes = onelaunch_step_bytes(pb, ndisp, args.mode, stored)
    else:
        dom = max(ms_k, key=ms_k.get)
        step_bytes = step_algorithmic_bytes(pb, ndisp, args.mode, stored)
    gbs = {k: nbytes[k] / (ms_k[k] / 1e3) / 1e9 if ms_k[k] > 0 else 0.0 for k in ms_k}
    step_gbs = step_bytes / (ms_per_step / 1e3) / 1e9
    probes = s.probes()
    finite = bool(np.isfinite(probes).all())
    s.close()

    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    # (scaled by a plain copy of known size) for this workload and mode
    traffic = None
    tf = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        if t.get("workload") == pb.name and t.get("mode") == args.mode:
            traffic = t.get("bytes_per_launch", {}).get(dom)

    batch = None
    if not args.no_batch:
        batch = batch_leg(args.batch_instances, args.batch_steps, args.plain_sample, args.mode)
    served = None
    if not args.no_serve:
        served = serve_leg(args.serve_requests, args.serve_rate, args.mode)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.mode == "fast" else "f32 (no FMA) + f64 poles",
        "data": "synthetic",
        "config": {
            "workload": f"{pb.name}: {pb.nx}^3 cells incl. 12-cell CPML, 8 lossy spheres, "
                        f"2-pole medium in 20% of the volume, TF/SF plane wave (+z, Ex), "
                        f"{len(pb.probes)} probes",
            "mode": args.mode,
            "cells": pb.cells,
            "dispersive_edges": ndisp,
            "tiling": tiling,
            "one_launch_step": onelaunch,
            "step_classes": classes,
            "step_algorithmic_bytes": step_bytes,
            "step_algorithmic_gbs": round(step_gbs, 1),
            "step_hbm_roofline_frac": round(step_gbs / HBM_PEAK_GBS, 4),
            "kernels": {k: {"ms": round(ms_k[k], 4), "algorithmic_bytes": nbytes[k],
                            "gbs": round(gbs[k], 1)} for k in ms_k},
            "probes_finite": finite,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs[dom], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs[dom] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": f"{KERNELS[dom]}, {nbytes[dom]} algorithmic B per launch",
        },
        "cpu_baseline": cpu,
        "batch": batch,
        "serve": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
