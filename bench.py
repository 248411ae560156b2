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
    value and per dispersive edge 8 B of Eprev plus the pole state (an upper
    bound: tiles that lie in one medium read no media indices)."""
    pe, ph = psi_counts(pb)
    per_state = 16 if mode == "fast" else 32
    return (onelaunch_cell_bytes(pb) * pb.cells + 8 * (pe + ph)
            + ndisp_edges * (8 + per_state * stored_poles))


_BITS = {0: "plain", 1: "cpml", 2: "poles", 3: "cpml+poles", 4: "sources",
         5: "cpml+sources", 6: "poles+sources", 7: "cpml+poles+sources"}
# tile classes of the one-launch step: 0-7 the k_step variants, 8 the
# four-cell interior kernel, 9-16 the half-width (32 lanes per row) variants
CLASS_NAMES = dict(_BITS)
CLASS_NAMES[8] = "plain4"
CLASS_NAMES.update({9 + c: "half:" + n for c, n in _BITS.items()})

KERNELS = {
    "e_face": "k_epass (Yee E + CPML E, fused face pass)",
    "ade": "k_ade_box (pole-residue ADE over the dispersive edges)",
    "h": "k_hpass (Yee H + CPML H, fused face pass)",
    "step_plain": "k_step<R, false, false, false, media type> (one launch per step, interior "
                  "tiles: Yee E and H of each plane)",
    "step_plain4": "k_step4<R, media type> (one launch per step, interior tiles four x cells "
                   "per lane: Yee E and H of each plane)",
}


def _cpu_child(n_sample, steps, variant):
    """Child process of cpu_baseline: on the heterogeneous workload at
    n_sample^3, the oracle (variant "nofma" or "fma") or the reference's own
    modules built with OpenMP (variant "reference"); prints one JSON line."""
    import oracle
    from semba_fdtd_amd import problem as P
    pb = P.heterogeneous_problem(n=n_sample, L=max(4, 12 * n_sample // 1024), nsteps=steps + 2,
                                 host_media=True)
    if variant == "reference":
        o = oracle.Reference(pb, omp=True)
    else:
        o = oracle.Oracle(pb, fma=variant == "fma")
    o.run(2)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    print(json.dumps({"cells": pb.cells, "steps": steps, "seconds": dt,
                      "L": pb.cpml_layers[0][0]}), flush=True)


def physical_cores():
    """Physical cores of the host (lscpu: cores per socket x sockets)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        kv = {k.strip(): v.strip() for k, _, v in (ln.partition(":") for ln in out.splitlines())}
        return int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(n_sample, steps):
    """CPU throughput on the host's physical cores -- at most the CPU share
    the process is given (OMP_NUM_THREADS) -- bound close, one thread per
    core, on the same heterogeneous workload at n_sample^3: the reference's
    own compiled modules (oracle/_ref_omp: calc_constants, borderscpml,
    electricdispersive, planewaves ... with their OpenMP directives, driven
    by the harness that restates the Yee loops of timestepping.F90) where
    that build travelled with the tree, and the oracle (C restatement, OpenMP
    over k/j) with FMA contraction off (its parity build) and on (an
    x86-64-v3 build).  Each runs in a child process so that the OpenMP
    placement applies."""
    import subprocess
    import oracle
    cores = physical_cores()
    threads = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    variants = ["nofma", "fma"] + (["reference"] if os.path.exists(oracle.REF_OMP_SO) else [])
    res = {}
    for v in variants:
        # the reference's serial pole loop makes it about 5x slower: a
        # quarter of the steps keeps its sample near the others' length
        ns = max(8, steps // 4) if v == "reference" else steps
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child",
                              str(n_sample), str(ns), v], env=env, capture_output=True,
                             text=True, timeout=900, check=True)
        res[v] = json.loads(out.stdout.strip().splitlines()[-1])
    rate = {v: r["cells"] * r["steps"] / r["seconds"] / 1e6 for v, r in res.items()}
    r = res["fma"]
    kind = "reference" if "reference" in rate else "port"
    value = rate["reference"] if kind == "reference" else rate["fma"]
    ref_txt = (f"reference modules (OpenMP build) {rate['reference']:.1f}; "
               if kind == "reference" else "")
    return {"value": round(value, 2), "unit": "Mcell-updates/s", "cores": threads, "kind": kind,
            "port_fma_on": round(rate["fma"], 2), "port_fma_off": round(rate["nofma"], 2),
            "physical_cores": cores,
            "sample": f"{threads} threads of {cores} physical cores, OMP_PROC_BIND=close "
                      f"OMP_PLACES=cores, heterogeneous workload at {n_sample}^3 (CPML "
                      f"{r['L']}, spheres, 20% 2-pole medium, complex128 poles), {steps} steps "
                      f"(reference: {max(8, steps // 4)}): "
                      f"{ref_txt}oracle fdtd_oracle.c -O2 FMA on (x86-64-v3) "
                      f"{rate['fma']:.1f}, FMA off {rate['nofma']:.1f} Mcell-updates/s"}


def batch_leg(count, steps, plain_sample, mode, storage="fp32"):
    """Batched uneven instances (log-uniform 32-128 cells per axis, 8-cell
    CPML, 1-6 blobs, plane wave): all instances advanced together, against a
    seeded random sample of the same instances solved one at a time, each on
    its own captured graph.  Both sides are set up and warmed (2 steps)
    before their timers start: the timed work is the same steps of the same
    kind of grid."""
    import torch
    import semba_fdtd_amd as sf
    from semba_fdtd_amd import problem as P
    from semba_fdtd_amd.batch import Batch
    insts = P.batch_instances(count=count, nsteps=steps)
    cells = sum(p.cells for p in insts)
    print(f"[bench] batch {storage}: {count} instances, {cells / 1e6:.1f} M cells, {steps} steps",
          file=sys.stderr, flush=True)
    sol = [sf.Solver(p, mode=mode, probe_buffer=steps, storage=storage) for p in insts]
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
    # plain path on a seeded random sample: each instance alone, set up and
    # warmed before its timer
    rng = np.random.default_rng(1004)
    pick = sorted(rng.choice(count, size=min(plain_sample, count), replace=False).tolist())
    scells, tp = 0, 0.0
    for q in pick:
        s = sf.Solver(insts[q], mode=mode, probe_buffer=steps, storage=storage)
        s.run(2)
        s.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(steps - 2)
        s.synchronize()
        tp += time.perf_counter() - t0
        scells += insts[q].cells
        s.close()
    plain = scells * (steps - 2) / tp / 1e6
    return {"instances": count, "cells": cells, "steps": steps, "mode": mode,
            "storage": storage,
            "batched_mcell_updates_per_s": round(batched, 1),
            "plain_mcell_updates_per_s": round(plain, 1),
            "plain_sample_instances": len(pick),
            "speedup": round(batched / plain, 2), "probes_finite": finite,
            "note": "batched: one persistent face-pass launch per half step for all instances, "
                    "graph-replayed; plain: a seeded random sample (numpy default_rng(1004)) "
                    "of the same instances, each alone on its own graph; both timed after "
                    "set-up and 2 warm-up steps"}


def serve_leg(count, rate, mode, storage="fp32"):
    """Serving stream (Poisson arrivals, cubic 48-192, 500-4000 steps, 30%
    with a 1-pole region, 64-frequency FP64 DFT): continuous batching against
    the plain one-solve-at-a-time path on the same stream."""
    from semba_fdtd_amd import problem as P
    from semba_fdtd_amd import serve
    reqs = P.serving_requests(count=count, rate=rate)
    print(f"[bench] serve {storage}: {count} requests at {rate}/s", file=sys.stderr, flush=True)
    plain = serve.serve_plain(reqs, mode=mode, storage=storage)
    # at most SERVE_MAX_CELLS cells admitted at once: enough to fill the
    # device, few enough that early requests finish early
    batched = serve.serve_batched(reqs, mode=mode, storage=storage, max_cells=SERVE_MAX_CELLS)
    same = True
    pm = {r.name: r for r in plain}
    for r in batched:
        same = same and np.array_equal(r.probes.view(np.uint32), pm[r.name].probes.view(np.uint32))
    bs, ps = serve.summarize(batched), serve.summarize(plain)
    return {"requests": count, "arrival_rate_per_s": rate, "mode": mode, "storage": storage,
            "batched_max_cells": SERVE_MAX_CELLS,
            "batched": bs, "plain": ps,
            "throughput_ratio": round(bs["mcell_updates_per_s"] / ps["mcell_updates_per_s"], 2),
            "results_identical": bool(same)}


def config_rows(cpu_seconds=20.0):
    """Per-configuration throughput beside the 1024^3 line (``--configs``):
    the plain-method cavity (64^3, PEC walls, Hertzian dipole, 2000 steps) and
    the open scatterer (256^3, 10-cell CPML, graded z, TF/SF wave, sphere,
    4000 steps), each run for its full step count on the GPU (FAST, set up and
    warmed before the timer; the autotuner picks the decomposition), and the
    oracle (OpenMP over the host's physical cores, bound close, FMA on and
    off) timed on a bounded number of steps of the same grid on the host.
    Algorithmic bytes: 84 B per cell plus 8 B per psi value per step."""
    import subprocess
    import torch
    import semba_fdtd_amd as sf
    from semba_fdtd_amd import problem as P
    from semba_fdtd_amd import tuning
    cores = physical_cores()
    threads = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    out = {}
    for name, pb in (("config1_cavity64", P.cavity_problem(n=64, nsteps=2000)),
                     ("config2_open_scatterer256", P.open_scatterer_problem(n=256, nsteps=4000))):
        ns = pb.nsteps_max
        print(f"[bench] {name}: {pb.nx}x{pb.ny}x{pb.nz}, {ns} steps", file=sys.stderr, flush=True)
        s = sf.Solver(pb, mode="fast", probe_buffer=ns)
        # the grid is at rest and the timed passes run no sources: it stays
        # at rest, so the sweep needs no checkpoint
        tuned = tuning.autotune(s, reps=3, keep_state=False)
        s.run(2)
        s.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(ns - 2)
        s.synchronize()
        dt = time.perf_counter() - t0
        finite = bool(np.isfinite(s.probes()).all())
        s.close()
        rate = pb.cells * (ns - 2) / dt / 1e6
        pe, ph = psi_counts(pb)
        bpc = 84.0 + 8.0 * (pe + ph) / pb.cells
        row = {"cells": pb.cells, "steps": ns, "gpu_mcell_updates_per_s": round(rate, 1),
               "gpu_seconds": round(dt, 3), "algorithmic_bytes_per_cell": round(bpc, 2),
               "hbm_roofline_frac_84B": round(rate * 1e6 * bpc / (HBM_PEAK_GBS * 1e9), 4),
               "decomposition": {"passes": list(tuned["tiling"]), "step": list(tuned["step"])},
               "probes_finite": finite}
        # the oracle on a bounded sample of the same grid's steps
        sample = max(4, int(cpu_seconds * 100e6 / pb.cells))
        for v in ("fma", "nofma"):
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-config",
                                name, str(min(sample, ns)), v], env=env, capture_output=True,
                               text=True, timeout=900, check=True)
            res = json.loads(r.stdout.strip().splitlines()[-1])
            row[f"cpu_oracle_{v}_mcell_updates_per_s"] = round(
                res["cells"] * res["steps"] / res["seconds"] / 1e6, 2)
        row["cpu_threads"] = threads
        row["cpu_sample_steps"] = min(sample, ns)
        out[name] = row
        print(f"[bench] {name}: {json.dumps(row)}", file=sys.stderr, flush=True)
    return out


def _config_problem(name, steps):
    """The configuration's grid for a CPU sample of `steps` steps."""
    from semba_fdtd_amd import problem as P
    if name.startswith("config1"):
        return P.cavity_problem(n=64, nsteps=steps + 2)
    if name.startswith("config2"):
        return P.open_scatterer_problem(n=256, nsteps=max(steps + 2, 200))
    if name.startswith("config3"):
        return P.heterogeneous_problem(n=256, L=4, nsteps=steps + 2, host_media=True)
    raise ValueError(name)


def _cpu_config(name, steps, variant):
    """Child process of config_rows and cpu_table: on the named configuration
    for `steps` steps, the oracle with FMA contraction on ("fma") or off
    ("nofma"), or the reference's own modules built with OpenMP
    ("reference"; one instance per process: their module state is global);
    prints one JSON line."""
    import oracle
    if name.startswith("config4"):
        # a seeded sample of 8 of the 256 batched instances, one after another
        from semba_fdtd_amd import problem as P
        insts = P.batch_instances(count=256, nsteps=1500)
        pick = sorted(np.random.default_rng(1004).choice(256, size=8, replace=False).tolist())
        cells, dt = 0, 0.0
        for q in pick:
            pb = insts[q]
            pb.nsteps_max = steps + 2
            o = oracle.Oracle(pb, fma=variant == "fma")
            o.run(2)
            t0 = time.perf_counter()
            o.run(steps)
            dt += time.perf_counter() - t0
            cells += pb.cells
        print(json.dumps({"cells": cells, "steps": steps, "seconds": dt, "instances": pick}),
              flush=True)
        return
    pb = _config_problem(name, steps)
    if variant == "reference":
        o = oracle.Reference(pb, omp=True)
    else:
        o = oracle.Oracle(pb, fma=variant == "fma")
    o.run(2)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    print(json.dumps({"cells": pb.cells, "steps": steps, "seconds": dt}), flush=True)


def cpu_table(threads):
    """CPU throughput per configuration (``--cpu-table``, no GPU needed): the
    oracle with FMA contraction on and off and, where oracle/_ref_omp was
    built, the reference's own modules with their OpenMP directives, each in
    a child process on `threads` threads bound close to cores, on bounded
    samples: 64^3 cavity 2000 steps (reference 500), 256^3 open scatterer 40
    steps (reference 10), the heterogeneous workload at 256^3 with a 4-cell
    CPML 40 steps (reference 10), 8 seeded batched instances 40 steps each
    (oracle only)."""
    import subprocess
    import oracle
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    plan = [("config1_cavity64", 2000, 500), ("config2_open_scatterer256", 40, 10),
            ("config3_heterogeneous256_sample", 40, 10), ("config4_batch_sample8", 40, 0)]
    out = {}
    for name, s1, s2 in plan:
        row = {}
        variants = [("fma", s1), ("nofma", s1)]
        if s2 and os.path.exists(oracle.REF_OMP_SO):
            variants.append(("reference", s2))
        for v, ns in variants:
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-config", name,
                                str(ns), v], env=env, capture_output=True, text=True,
                               timeout=1800, check=True)
            res = json.loads(r.stdout.strip().splitlines()[-1])
            row[v] = round(res["cells"] * res["steps"] / res["seconds"] / 1e6, 2)
            row[f"{v}_steps"] = ns
        out[name] = row
        print(f"[bench] {name}: {json.dumps(row)}", file=sys.stderr, flush=True)
    return {"threads": threads, "unit": "Mcell-updates/s", "rows": out}


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
    ap.add_argument("--cpu-steps", type=int, default=100)
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
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--configs", action="store_true",
                    help="also time the plain-method cavity (64^3) and the open scatterer "
                         "(256^3) at their full step counts, with the oracle on the host")
    if len(sys.argv) == 5 and sys.argv[1] == "--cpu-child":
        _cpu_child(int(sys.argv[2]), int(sys.argv[3]), sys.argv[4])
        return
    if len(sys.argv) == 5 and sys.argv[1] == "--cpu-config":
        _cpu_config(sys.argv[2], int(sys.argv[3]), sys.argv[4])
        return
    if len(sys.argv) >= 2 and sys.argv[1] == "--cpu-table":
        # CPU only: no GPU is touched
        threads = int(sys.argv[2]) if len(sys.argv) > 2 else physical_cores()
        print(json.dumps(cpu_table(threads)), flush=True)
        return
    args = ap.parse_args()
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
        # one launch per step: each tile class's kernel timed alone; its
        # compulsory bytes are onelaunch_cell_bytes per owned cell plus the
        # CPML psi and pole state of those cells (the library counts them from
        # its tile plan)
        ms_k = {}
        classes = []
        cb = onelaunch_cell_bytes(pb)
        timed_classes = s.step_classes(reps)
        cbytes = s.step_class_bytes()
        for cls, cells, ms_c in timed_classes:
            key = "step_" + CLASS_NAMES[cls].replace("+", "_").replace(":", "_")
            ms_k[key] = ms_c
            nbytes[key] = int(cbytes[cls])
            classes.append({"class": CLASS_NAMES[cls], "cells": int(cells), "ms": round(ms_c, 4),
                            "algorithmic_bytes": int(cbytes[cls]),
                            "gbs": round(cbytes[cls] / (ms_c / 1e3) / 1e9, 1)})
        dom = max(ms_k, key=ms_k.get)
        KERNELS["step_plain"] += f", {cb} B per owned cell"
        KERNELS["step_plain4"] += (f", {cb} B per owned cell, 48 B in single-medium tiles "
                                   "(no media indices read)")
        if dom not in ("step_plain", "step_plain4"):
            KERNELS[dom] = (f"k_step tile class {dom[5:]} ({cb} B per owned cell, 8 B per psi "
                            f"value, 16 B per stored pole per pole-state cell, + 8 B of "
                            f"previous field in tiles with sources)")
        # the classes' own compulsory bytes (single-medium tiles read no media)
        step_bytes = int(sum(cbytes[c] for c, _, _ in timed_classes))
    else:
        dom = max(ms_k, key=ms_k.get)
        step_bytes = step_algorithmic_bytes(pb, ndisp, args.mode, stored)
    gbs = {k: nbytes[k] / (ms_k[k] / 1e3) / 1e9 if ms_k[k] > 0 else 0.0 for k in ms_k}
    step_gbs = step_bytes / (ms_per_step / 1e3) / 1e9
    # the whole step against the half-step yardstick (84 B per cell, 8 B per
    # psi value, 8 + 16 B per stored pole per dispersive edge) and against the
    # one-launch step's own compulsory bytes
    bytes84 = step_algorithmic_bytes(pb, ndisp, args.mode, stored)
    step_roofline = {
        "yardstick_84B": {"bytes": bytes84,
                          "gbs": round(bytes84 / (ms_per_step / 1e3) / 1e9, 1),
                          "frac": round(bytes84 / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "onelaunch": {"bytes": step_bytes, "gbs": round(step_gbs, 1),
                      "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
    }
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

    batch = batch16 = None
    if not args.no_batch:
        batch = batch_leg(args.batch_instances, args.batch_steps, args.plain_sample, args.mode)
        if args.mode == "fast":
            batch16 = batch_leg(args.batch_instances, args.batch_steps, args.plain_sample,
                                args.mode, storage="fp16")
    served = served16 = None
    if not args.no_serve:
        served = serve_leg(args.serve_requests, args.serve_rate, args.mode)
        if args.mode == "fast":
            served16 = serve_leg(args.serve_requests, args.serve_rate, args.mode, storage="fp16")

    configs = config_rows() if args.configs else None

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
            "step_roofline": step_roofline,
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
        "batch_fp16": batch16,
        "serve": served,
        "serve_fp16": served16,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
