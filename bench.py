#!/usr/bin/env python3
"""Throughput of the fused FDTD time step on one MI355X.

A "step" is one full FDTD time step of the grid: the fused H/D/source/Debye-E
kernel, the Mur face kernel (which also records the probe row) and (every
500 steps) the z-slice capture.  The default workload is the 320x320x880
voxel phantom of 50 Debye tissues with a 10-cell Jz dipole at 900 MHz, 3
probes and a z-slice every 500 steps (large_phantom in
fdtd_amd/workloads.py).  The grid first advances --advance steps untimed (so
the timed steps run on a populated field), then --warmup steps, then the
--steps timed steps; the line states the timed step range.

Prints one JSON line: Gcell-updates/s over the timed steps (HIP events on
the bench's stream around the whole step sequence), the fused kernel's
roofline (97 algorithmic bytes per cell-update over its mean launch
duration, from HIP events bound to each dispatch over --timing-launches
further steps, at least 64, taken after the timed steps so timing costs the
timed steps nothing) and the CPU baseline (the oracle's FP32 transcription,
OpenMP on this host's cores, on a bounded sample of the same workload).
Exits with status 3 when the probes saw no finite non-zero signal.

Two more workloads measure the systems patterns:
  --workload batched  64 uneven grids (batched_set) stepped together, one
                      persistent fused launch per step from captured graphs,
                      shape picked by the autotuner (or fixed by --shape
                      64xTYxKC); whole-job Gcell-updates/s
                      beside running the same grids one after another.
  --workload riken    the RIKEN voxel phantom at body scale (297x522x1719 with
                      its vacuum pad), from `--phantom` (a PGM directory or
                      ids packed by `python -m fdtd_amd.phantom`).
  --workload served   an open-loop Poisson stream of solves (served_stream)
                      through the continuous-batching server: solves/s and
                      p50/p99 latency beside the sequential drop-in path (the
                      five reference entry points, one solve at a time, FIFO)
                      on the same stream.
  --workload stream   the memory ceiling of the fused step's access mix: 15
                      arrays read and 9 written per cell with no arithmetic
                      (fdtd_stream_probe), in GB/s.
  --workload cli      the `fdtd` command-line driver on the reference's own
                      parameter file (data/input_params: 100^3, 3 GHz, 1 mm
                      cells) with synthetic tissue slices, --steps steps and all
                      nine fields dumped as text after every step: wall time of
                      the overlapped result pipeline beside --sync (each dump
                      written on the stepping thread).
  --workload slabs    320x320x880 cut into --slabs z-slabs (each with ghost
                      planes), all stepped on this one GPU with the halo
                      exchange of every interface before every step: the
                      part of a multi-GPU z-slab run one GPU can show, beside
                      the unsplit grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CELL = {"f32": 97, "f16": 49, "bf16": 49}  # 24 state words + 1 material byte per cell-update
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--advance", type=int, default=300,
                    help="untimed steps before the warm-up, so the timed steps run on a populated field")
    ap.add_argument("--timing-launches", type=int, default=64,
                    help="fused launches timed one by one (HIP events bound to each dispatch) after the timed "
                         "steps, for roofline.achieved; at least 64")
    ap.add_argument("--workload", default="large_phantom",
                    choices=["plain_256", "large_phantom", "plain_small", "batched", "served", "riken", "slabs",
                             "stream", "cli"])
    ap.add_argument("--slabs", type=int, default=8, help="slabs: z-slabs of the 320x320x880 grid")
    ap.add_argument("--grid", default=None,
                    help="slabs: NXxNYxNZ of a box phantom in place of the 320x320x880 grid (any size that "
                         "fits the device as slabs, e.g. 1024x1024x2500)")
    ap.add_argument("--phantom", default=os.path.join(ROOT, "data", "riken_ids.npz"),
                    help="riken: PGM slice directory or ids packed by `python -m fdtd_amd.phantom`")
    ap.add_argument("--specs", default=os.path.join(ROOT, "tests", "golden", "mat_specs_riken"),
                    help="riken: tissue table")
    ap.add_argument("--requests", type=int, default=120, help="served: requests of the stream")
    ap.add_argument("--f16-share", type=float, default=0.5,
                    help="served: fraction of requests with fp16 storage (0: every request fp32, within "
                         "rel-L2 1e-4 of the FP64 oracle)")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="served: device memory budget")
    ap.add_argument("--slots", type=int, default=64, help="served: grids on the device at once")
    ap.add_argument("--segment", type=int, default=8, help="batched/served: steps per graph replay")
    ap.add_argument("--no-sequential", action="store_true", help="batched/served: skip the sequential baselines")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"],
                    help="field storage (f32 reproduces the reference's arithmetic bit for bit)")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="steps of the CPU baseline sample (0: about 5 G cell-updates, 10-15 s of host work on "
                         "16 threads)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch shape by timing each candidate first (default: the library's "
                         "default shape for the grid, 64x4 tiles of 16 planes at 320x320x880, so the "
                         "command's profile is reproducible)")
    ap.add_argument("--no-autotune", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--load-zero-terms", action="store_true",
                    help="fp32: load E^{n-1} and D^{n-1} on every row, also where the Debye update multiplies "
                         "them by zero (tau = 0); the default skips those loads, bit for bit the same results")
    ap.add_argument("--xcd-remap", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="block -> tile order of the fused kernel (FdtdExt.xcd_remap; default: the library's)")
    ap.add_argument("--time-steps", type=int, default=1, choices=[1, 2],
                    help="fp32: time steps per fused launch (2: the two-step kernel where it applies)")
    ap.add_argument("--shape", default=None,
                    help="fixed launch shape TXxTYxKC (threads per tile row, tile rows, z planes per block), "
                         "e.g. 128x2x6; no autotuning (batched: 64xTYxKC)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run one device-to-device copy of this many bytes first (HBM counter calibration)")
    ap.add_argument("--traffic-kib", type=float, nargs=2, default=None,
                    help="measured FETCH_SIZE and WRITE_SIZE per fused launch (KiB, already scaled)")
    return ap.parse_args()


def build_workload(name, steps, a=None):
    from fdtd_amd import workloads as W
    if name == "riken":
        import fdtd_amd as F
        from fdtd_amd import phantom as P
        table, _ = F.load_tissue_table(a.specs)
        return W.riken(P.load(a.phantom, P.RIKEN_SHAPE), table, steps=steps)
    if name == "plain_256":
        return W.plain_256(steps=steps)
    if name == "large_phantom":
        return W.large_phantom(steps=steps)
    return W.plain_small(steps=steps)


PROFILES = [os.path.join(ROOT, "profiles", f) for f in ("fused_step_large_phantom.md", "fused_step_plain_256.md",
                                                          "fused_step_plain256_large_phantom.md",
                                                          "fused_step_riken.md", "fused_step_large_phantom_f16.md")]


def profiled_traffic(workload, storage, launch, loading=False, order=0):
    """(HBM bytes per fused launch, profile path) measured for this workload
    at this launch shape and block order (rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE passes, scaled by a 1 GiB copy) in a committed profile, or
    (None, None).  Sections are headed "## <workload> at <shape>, <order>
    order[, <storage> storage][, every row loading]..."; the kernel that
    loads E^{n-1} and D^{n-1} on every row (loading=True) has its own
    sections."""
    import re
    names = {0: "identity order", 3: "row order"}
    if order not in names:
        return None, None
    for path in PROFILES:
        if not os.path.exists(path):
            continue
        for sec in open(path).read().split("\n## ")[1:]:
            head = sec.split("\n", 1)[0]
            m = re.match(re.escape(f"{workload} at {launch}") + r"(?![\dx])(.*)", head)
            if not m or names[order] not in m.group(1) or ("every row loading" in m.group(1)) != loading:
                continue
            # fp32 sections name no storage; 16-bit ones say "<storage> storage"
            if (storage != "f32") != ("storage" in m.group(1)) or (storage != "f32" and f"{storage} storage" not in m.group(1)):
                continue
            rd = re.search(r"measured HBM read \(scaled FETCH_SIZE\) \| [\d.]+ GB \((\d+) KiB x (\d+)\)", sec)
            wr = re.search(r"measured HBM write \(scaled WRITE_SIZE\) \| [\d.]+ GB \((\d+) KiB\)", sec)
            if rd and wr:
                return (int(rd.group(1)) * int(rd.group(2)) + int(wr.group(1))) * 1024, path
    return None, None


def zero_term_fraction(w, shape):
    """Fraction of the fused kernel's (32-cell half-row segment, plane) units
    that skip loading E^{n-1} and D^{n-1}: every cell of the segment (half a
    wave, one 128-byte line of each array) has tau = 0 (exactly +0), and the
    kernel has decided for the plane (a z-chunk's first plane always loads,
    and the bottom chunk decides from its third plane)."""
    nx, ny, nz = w.shape
    kc = shape["kchunk"]
    zero_mat = np.asarray([np.float32(t[3]).view(np.uint32) == 0 for t in w.table])
    zc = zero_mat[w.ids]                                   # (nz, ny, nx)
    segs = (nx + 31) // 32
    pad = np.ones((nz, ny, segs * 32), bool)
    pad[:, :, :nx] = zc
    seg_ok = pad.reshape(nz, ny, segs, 32).all(axis=3)     # (nz, ny, segs)
    k = np.arange(nz)
    decided = ((k % kc) != 0) & ~((k < kc) & (k == 1))
    return float((seg_ok & decided[:, None, None]).sum() / seg_ok.size)


def host_cpu():
    """The host's CPU model name, from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, steps):
    """The oracle's FP32 transcription on this host's cores (a baseline only):
    `steps` steps on up to 16 threads, then steps/25 on one thread."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    nx, ny, nz = w.shape

    def timed(nthreads, nsteps):
        O.set_threads(nthreads)
        sim = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0,
                    prec=32)
        sim.set_materials(*w.material_arrays())
        sim.set_sources(w.sources, w.jz, jz_mode=1)
        sim.step(1)  # first touch of the arrays
        t0 = time.perf_counter()
        sim.step(nsteps)
        return time.perf_counter() - t0

    dt = timed(threads, steps)
    steps1 = max(2, steps // 25)
    dt1 = timed(1, steps1)
    return {"value": round(nx * ny * nz * steps / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{w.name} {nx}x{ny}x{nz}, {steps} steps of the oracle FP32 transcription "
                      f"(OpenMP over z, {threads} threads), {dt:.2f} s",
            "single_core": {"value": round(nx * ny * nz * steps1 / dt1 / 1e9, 4), "cores": 1,
                            "sample": f"{steps1} steps on one thread, {dt1:.2f} s"},
            "host": {"cpu": host_cpu(), "nproc": os.cpu_count()}}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not supported",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import fdtd_amd as F
    from fdtd_amd import hip

    if a.copy_bytes:
        src, dst = hip.DeviceArray(a.copy_bytes // 4), hip.DeviceArray(a.copy_bytes // 4)
        hip.ok(hip.rt().hipMemcpy(dst.ptr, src.ptr, a.copy_bytes, hip.D2D), "calibration copy")
        hip.sync()
        src.free()
        dst.free()

    if a.workload == "batched":
        return bench_batched(a)
    if a.workload == "served":
        return bench_served(a)
    if a.workload == "slabs":
        return bench_slabs(a)
    if a.workload == "stream":
        return bench_stream(a)
    if a.workload == "cli":
        return bench_cli(a)

    timing = max(64, a.timing_launches)
    total = a.advance + a.warmup + a.steps + timing
    w = build_workload(a.workload, total, a)
    nx, ny, nz = w.shape
    cells = nx * ny * nz

    stream = hip.Stream()
    storage = {"f32": F.STORE_F32, "f16": F.STORE_F16, "bf16": F.STORE_BF16}[a.storage]
    bpc = BYTES_PER_CELL[a.storage]
    fixed = {}
    if a.shape:
        tx, ty, kc = (int(v) for v in a.shape.lower().split("x"))
        fixed = dict(tile_x=tx, tile_y=ty, kchunk=kc, lane_cells=1)
    ctx = F.Context(w.params, storage=storage, load_zero_terms=int(a.load_zero_terms), time_steps=a.time_steps,
                    xcd_remap=a.xcd_remap, **fixed)
    remap = a.xcd_remap if a.xcd_remap is not None else F.make_ext().xcd_remap
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, total)
    if w.slice_every:
        ctx.set_slices(w.slice_k, w.slice_every, total // w.slice_every + 1)
    tuned = None
    if a.autotune and not a.shape:
        # untimed: the autotuner runs each launch shape a few steps, then resets the grid to step 0
        tuned = ctx.autotune(trial_steps=8, stream=stream.handle)
    ctx.run(a.advance + a.warmup, stream.handle)
    stream.sync()
    first = ctx.n

    e0, e1 = hip.Event(), hip.Event()
    hip.sync()
    e0.record(stream)
    ctx.run(a.steps, stream.handle)
    e1.record(stream)
    e1.sync()
    hip.sync()
    ms = e0.elapsed_ms(e1)
    # the fused kernel's own duration: every launch of `timing` further steps, events bound to the dispatch
    ctx.set_timing(timing, every=1)
    ctx.run(timing, stream.handle)
    stream.sync()
    kms, launches = ctx.kernel_time()
    per_launch = max(1, ctx.timed_steps())
    probes = ctx.probes()
    ok = bool(np.isfinite(probes).all()) and bool(np.abs(probes).max() > 0)

    value = cells * a.steps / (ms * 1e-3) / 1e9
    kavg_s = kms / launches * 1e-3
    achieved = bpc * cells * per_launch / kavg_s / 1e9
    shape = ctx.launch_shape()
    launch = f"{shape['tile_x']}x{shape['tile_y']}x{shape['kchunk']}"
    traffic, traffic_src = None, None
    if a.traffic_kib:
        traffic, traffic_src = int((a.traffic_kib[0] + a.traffic_kib[1]) * 1024), "--traffic-kib"
    elif per_launch != 1:
        traffic, traffic_src = None, "no counter profile of the two-step kernel under profiles/"
    else:
        traffic, prof = profiled_traffic(w.name, a.storage, launch, bool(a.load_zero_terms), remap)
        traffic_src = (os.path.relpath(prof, ROOT) + f" (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE, 1 GiB copy "
                       f"scale; {launch} launch, block order {remap})") if traffic is not None else \
            f"no counter profile of {w.name} at launch {launch}, block order {remap}" + \
            (" (every row loading)" if a.load_zero_terms else "") + " under profiles/"
    out = {
        "metric": "Gcell-updates/s (fused FDTD step, Debye media, Mur ABC)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 Debye sums)" if a.storage == "f32" else f"{a.storage} storage, f32 arithmetic",
        "data": ("RIKEN voxel phantom (the reference's data/pgmdata), vacuum-padded" if w.name == "riken"
                 else "synthetic (seeded phantom)"),
        "config": {"workload": w.name, "grid": [nx, ny, nz], "cells": cells, "materials": int(w.table.shape[0]),
                   "sources": int(w.sources.shape[0]), "probes": int(w.probes.shape[0]),
                   "slice_every": w.slice_every, "launch": shape, "xcd_remap": remap, "time_steps": a.time_steps,
                   "autotuned": tuned is not None,
                   "timed_steps": [first, first + a.steps], "untimed_advance": a.advance,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "fused_step" if per_launch == 1 else "fused_step2",
                     "steps_per_launch": per_launch, "kernel_avg_ms": round(kms / launches, 5),
                     "launches_timed": launches,
                     "timed_launch_steps": [first + a.steps, first + a.steps + launches * per_launch],
                     "bytes_per_cell": bpc},
        "probes_finite_nonzero": ok,
    }
    if a.storage == "f32":
        # rows of the fused kernel that skip E^{n-1}, D^{n-1} (tau = 0 on all their cells): the bytes the
        # kernel needs per cell beside the 97 B/cell model
        zf = zero_term_fraction(w, shape)
        out["roofline"]["zero_term_rows"] = round(zf, 4) if not a.load_zero_terms else 0.0
        need = bpc - 24.0 * (zf if not a.load_zero_terms else 0.0)
        out["roofline"]["bytes_per_cell_needed"] = round(need, 2)
        out["roofline"]["achieved_needed_bytes"] = round(need * cells / kavg_s / 1e9, 1)
    ctx.close()
    # the same box's streaming ceiling for this access mix (fdtd_stream_probe), a yardstick beside the
    # peak.  16-bit fields move as 32-bit words of two cells, so their probe streams half as many words.
    # The probe's best launch shape depends on the footprint (resident blocks grid-striding at 1-2 GB,
    # 8192 or more blocks at 8 GB), so the ceiling is the best of three.
    words = cells if a.storage == "f32" else (cells + 1) // 2
    ceil_gbs = max(F.stream_probe(words, blocks, reps=10)[0] for blocks in (0, 8192, 16384))
    out["roofline"]["stream_ceiling"] = {
        "gb_per_s": round(ceil_gbs, 1), "frac": round(achieved / ceil_gbs, 4),
        "what": "15 arrays read + 9 written, no arithmetic, " + (
            "one fp32 word per cell" if a.storage == "f32" else f"{words} 32-bit words (two 16-bit cells each)")
        + "; best of the resident grid, 8192 and 16384 blocks"}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_steps or max(2, int(round(5e9 / cells))))
    print(json.dumps(out), flush=True)
    if not ok:
        print("bench: the probes saw no finite non-zero signal", file=sys.stderr, flush=True)
        sys.exit(3)


def bench_cli(a):
    """The CLI with a dump of all nine fields after every step, overlapped
    output against --sync, on the reference's 100^3 parameter file."""
    import re
    import shutil
    import subprocess
    import tempfile
    cli = os.path.join(ROOT, "fdtd_amd", "fdtd")
    specs = os.path.join(ROOT, "tests", "golden", "mat_specs_riken")
    listed = [int(l.split()[0]) for l in open(specs) if l[:1].isdigit()]
    steps = a.steps if a.steps % 3 == 0 else a.steps + 3 - a.steps % 3   # the parameter file rounds up to 3
    work = tempfile.mkdtemp(prefix="fdtd_cli_")
    try:
        pgm = os.path.join(work, "pgm")
        os.mkdir(pgm)
        rng = np.random.default_rng(12352)
        for k in range(100):
            vals = rng.choice(listed, 100 * 100)
            with open(os.path.join(pgm, f"v1_{k + 1:05d}.pgm"), "w") as f:
                f.write("P2\n# test\n100 100\n94\n" + "\n".join(str(int(v)) for v in vals) + "\n")
        text = open(os.path.join(ROOT, "tests", "golden", "input_params")).read()
        text = text.replace("t_max: 3", f"t_max: {steps}").replace("env_set_dir: data/pgmdata",
                                                                   f"env_set_dir: {pgm}")
        runs = {}
        for mode, extra in (("sync", ["--sync"]), ("overlapped", [])):
            out = os.path.join(work, "out_" + mode)
            os.mkdir(out)
            params = os.path.join(work, "params_" + mode)
            with open(params, "w") as f:
                f.write(text.replace("output_dir: output", f"output_dir: {out}"))
            t0 = time.perf_counter()
            r = subprocess.run([cli, params, "--specs", specs, "--quiet", "--every", "1", *extra],
                               capture_output=True, text=True, env=dict(os.environ, FDTD_TIMING="1"), timeout=1200)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(f"fdtd {mode}: {r.stdout[-2000:]}")
            m = re.search(r"Stepping and output: ([\d.]+) s; the stepping thread waited ([\d.]+) s", r.stdout)
            files = os.listdir(out)
            runs[mode] = {"process_s": round(wall, 3), "stepping_and_output_s": float(m.group(1)),
                          "stepping_thread_waited_s": float(m.group(2)), "files": len(files),
                          "bytes": sum(os.path.getsize(os.path.join(out, n)) for n in files)}
            print(f"[cli] {mode}: {runs[mode]}", file=sys.stderr, flush=True)
        same = all(open(os.path.join(work, "out_sync", n), "rb").read() ==
                   open(os.path.join(work, "out_overlapped", n), "rb").read()
                   for n in sorted(os.listdir(os.path.join(work, "out_sync"))))
    finally:
        shutil.rmtree(work, ignore_errors=True)
    t_sync, t_ovl = runs["sync"]["stepping_and_output_s"], runs["overlapped"]["stepping_and_output_s"]
    out = {
        "metric": "dumps/s (CLI, all nine fields as text after every step)",
        "value": round(steps / t_ovl, 3), "unit": "dumps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(t_ovl / steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic tissue slices (codes of data/mat_specs_riken)",
        "config": {"workload": "cli_input_params", "grid": [100, 100, 100], "every": 1, "writers": 7,
                   "parallelism": "single GPU"},
        "runs": runs, "speedup_vs_sync": round(t_sync / t_ovl, 3), "outputs_identical": same,
    }
    print(json.dumps(out), flush=True)


def bench_stream(a):
    """The memory ceiling of the fused step's access mix: 15 arrays read and
    9 written per cell, no arithmetic, every access nontemporal
    (fdtd_stream_probe_layout), at 256^3 and 320x320x880 cells, in two
    layouts (24 separate arrays as a context holds them, and row records:
    the 24 rows of one grid row contiguous), each at the resident block count
    and two larger grids."""
    import fdtd_amd as F
    reps = max(2, a.steps // 100)
    runs = {}
    for name, cells, row in (("256^3", 256 ** 3, 256), ("320x320x880", 320 * 320 * 880, 320)):
        for layout, w in (("arrays", 0), ("row_records", row)):
            for blocks in (0, 8192, 16384):
                gbs, ms = F.stream_probe_layout(cells, w, blocks, reps=reps)
                runs[f"{name} {layout} {'resident' if blocks == 0 else blocks}"] = {
                    "gb_per_s": round(gbs, 1), "ms_per_pass": round(ms, 5)}
                print(f"[stream] {name} {layout} blocks={blocks}: {gbs:.1f} GB/s", file=sys.stderr, flush=True)
    best = max(v["gb_per_s"] for v in runs.values())
    out = {
        "metric": "GB/s streaming the fused step's 15-read / 9-write access mix without arithmetic",
        "value": best, "unit": "GB/s", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (zeros)",
        "config": {"workload": "stream_mix", "arrays": {"read": 15, "written": 9}, "parallelism": "single GPU"},
        "by_layout": runs, "frac_of_peak": round(best / HBM_PEAK_GBS, 4),
    }
    print(json.dumps(out), flush=True)


def bench_slabs(a):
    """large_phantom (or a box phantom of `a.grid`) as `a.slabs` z-slabs on
    one GPU, halo exchange every step."""
    import fdtd_amd as F
    from fdtd_amd import hip
    from fdtd_amd import workloads as W
    total = a.warmup + a.steps
    if a.grid:
        w = W.box_phantom(tuple(int(v) for v in a.grid.lower().split("x")), steps=total)
    else:
        w = W.large_phantom(steps=total)
    nx, ny, nz = w.shape
    p = w.params
    n = max(2, a.slabs)
    bounds = [nz * s // n for s in range(n + 1)]
    stream = hip.Stream()
    s_ = stream.handle
    slabs = []
    free0 = hip.mem_info()[0]
    for s in range(n):
        r0, r1 = bounds[s], bounds[s + 1]
        g0, g1 = r0 - (s > 0), r1 + (s < n - 1)
        c = F.Context(F.make_params((nx, ny, g1 - g0), p.dt, p.dx, mu0=p.mu0, eps0=p.eps0))
        c.set_materials(w.ids[g0:g1], w.table)
        own = w.sources[(w.sources[:, 2] >= r0) & (w.sources[:, 2] < r1)].copy()
        own[:, 2] -= g0
        c.set_sources(own, w.jz)
        slabs.append(c)
    slab_bytes = free0 - hip.mem_info()[0]
    up_b, down_b = slabs[0].halo_bytes(F.HALO_UP), slabs[0].halo_bytes(F.HALO_DOWN)
    ups = [hip.DeviceArray(up_b // 4) for _ in range(n - 1)]
    downs = [hip.DeviceArray(down_b // 4) for _ in range(n - 1)]

    def exchange():
        for q in range(n - 1):
            lo, hi = slabs[q], slabs[q + 1]
            lo.halo_pack(F.HALO_UP, ups[q].ptr, s_)
            hi.halo_unpack(F.HALO_UP, ups[q].ptr, s_)
            hi.halo_pack(F.HALO_DOWN, downs[q].ptr, s_)
            lo.halo_unpack(F.HALO_DOWN, downs[q].ptr, s_)

    def step_all(k):
        for _ in range(k):
            exchange()
            for c in slabs:
                c.step(s_)

    step_all(a.warmup)
    stream.sync()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    step_all(a.steps)
    e1.record(stream)
    e1.sync()
    ms = e0.elapsed_ms(e1)
    # the exchange alone, same number of steps (the fields are left as they are)
    x0, x1 = hip.Event(), hip.Event()
    x0.record(stream)
    for _ in range(a.steps):
        exchange()
    x1.record(stream)
    x1.sync()
    xms = x0.elapsed_ms(x1)
    for c in slabs:
        c.close()
    for b in ups + downs:
        b.free()
    # the unsplit grid, same steps, where one context can hold it
    ums = None
    try:
        whole = F.Context(p)
    except F.FdtdError:
        whole = None
    if whole is not None:
        whole.set_materials(w.ids, w.table)
        whole.set_sources(w.sources, w.jz)
        whole.run(a.warmup, s_)
        stream.sync()
        u0, u1 = hip.Event(), hip.Event()
        u0.record(stream)
        whole.run(a.steps, s_)
        u1.record(stream)
        u1.sync()
        ums = u0.elapsed_ms(u1)
        whole.close()
    cells = nx * ny * nz
    out = {
        "metric": f"Gcell-updates/s ({nx}x{ny}x{nz} as {n} z-slabs on one GPU, halo exchange every step)",
        "value": round(cells * a.steps / (ms * 1e-3) / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms / a.steps, 5),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 (f64 Debye sums)",
        "data": "synthetic (seeded phantom)",
        "config": {"workload": f"{w.name}_slabs", "grid": [nx, ny, nz], "slabs": n,
                   "slab_planes": [bounds[s + 1] - bounds[s] for s in range(n)], "parallelism": "single GPU"},
        "exchange_us_per_step": round(xms / a.steps * 1e3, 2),
        "message_bytes_per_interface": {"up": up_b, "down": down_b},
        "unsplit_value": round(cells * a.steps / (ums * 1e-3) / 1e9, 3) if ums else None,
        "device_gb_of_slabs": round(slab_bytes / 1e9, 1),
    }
    print(json.dumps(out), flush=True)


def _instance(F, w, storage="f32", load_zero_terms=0):
    ext = F.make_ext(storage={"f32": F.STORE_F32, "f16": F.STORE_F16, "bf16": F.STORE_BF16}[storage],
                     load_zero_terms=load_zero_terms)
    return F.Instance(w.params, ext, w.ids, w.table, w.sources, w.jz, w.steps, w.probes)


def _pct(v, q):
    v = sorted(v)
    k = min(max(int(-(-q * len(v) // 1)), 1), len(v))
    return v[k - 1]


def _sequential_fused(F, hip, insts):
    """Each grid alone in a fused context, one after another (set-up, run,
    probe read-back and free per grid)."""
    times = []
    for inst in insts:
        t0 = time.perf_counter()
        e = inst.ext
        ctx = F.Context(inst.params, debye=e.debye_mode, jz_mode=e.jz_index_mode, storage=e.storage)
        ctx.set_materials(inst.ids, inst.table)
        ctx.set_sources(inst.src, inst.jz)
        ctx.set_probes(inst.probes, inst.steps)
        ctx.run(inst.steps)
        ctx.probes()
        ctx.close()
        times.append(time.perf_counter() - t0)
    return times


def _sequential_dropin(F, insts):
    """Each grid through the five drop-in entry points, one after another
    (fp32, the reference's 512-value jz table)."""
    times = []
    for inst in insts:
        ext = F.make_ext(jz_limit=512)
        x = F.Instance(inst.params, ext, inst.ids, inst.table, inst.src, inst.jz, inst.steps, inst.probes)
        t0 = time.perf_counter()
        F.dropin_solve(x)
        times.append(time.perf_counter() - t0)
    return times


def bench_batched(a):
    import fdtd_amd as F
    from fdtd_amd import hip
    from fdtd_amd import workloads as W

    fixed = None
    if a.shape:
        # a fixed shape (counter passes need the same kernel in every pass): 64 x tile rows x planes
        tx, fty, fkc = (int(v) for v in a.shape.lower().split("x"))
        if tx != 64:
            raise SystemExit("batched: --shape takes 64xTYxKC (64-wide tiles)")
        fixed = (fty, fkc)
    ws = W.batched_set()
    insts = [_instance(F, w, a.storage, int(a.load_zero_terms)) for w in ws]
    work = sum(w.cells * w.steps for w in ws)
    cells = sum(w.cells for w in ws)
    arena = int(sum(x.nbytes() for x in insts) * 1.02) + (64 << 20)
    b = F.Batch(arena)
    b.set_shape(graph_steps=a.segment)
    t0 = time.perf_counter()
    ids = [b.add(x) for x in insts]
    t_upload = time.perf_counter() - t0
    stream = hip.Stream()
    t0 = time.perf_counter()
    if fixed:
        ty, kc = fixed
        b.set_shape(tile_y=ty, kchunk=kc)
    else:
        ty, kc = b.autotune(trial_steps=4, stream=stream.handle)
    t_tune = time.perf_counter() - t0
    info = b.info()
    steps = max(w.steps for w in ws)
    e0, e1 = hip.Event(), hip.Event()
    hip.sync()
    e0.record(stream)
    b.run(steps, stream.handle)
    e1.record(stream)
    e1.sync()
    ms = e0.elapsed_ms(e1)
    done = [b.status(i)[0] == x.steps for i, x in zip(ids, insts)]
    series = [b.probes(i) for i in ids]
    ok = all(done) and all(np.isfinite(s).all() for s in series)
    b.close()
    value = work / (ms * 1e-3) / 1e9
    out = {
        "metric": "Gcell-updates/s (64 uneven grids batched: one persistent fused launch per step)",
        "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(ms / steps, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 Debye sums)" if a.storage == "f32" else f"{a.storage} storage, f32 arithmetic",
        "data": "synthetic (seeded phantoms)",
        "config": {"workload": "batched_set", "grids": len(ws), "cells": cells, "cell_updates": work,
                   "steps_range": [min(w.steps for w in ws), steps], "tile_y": ty, "kchunk": kc,
                   "graph_steps": a.segment, "persistent_grid": info["grid"][0], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(97 * work / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(97 * work / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "whole batch step (fused_batch + mur_batch with probe rows)",
                     "bytes_per_cell": 97},
        "upload_s": round(t_upload, 3), "autotune_s": round(t_tune, 3), "all_finished": ok,
    }
    # end to end, as the sequential runs below are timed: uploads, the autotuner and the run
    e2e = t_upload + t_tune + ms * 1e-3
    out["end_to_end"] = {"seconds": round(e2e, 3), "gcell_per_s": round(work / e2e / 1e9, 3)}
    if a.storage != "f32":
        out["roofline"]["bytes_per_cell"] = BYTES_PER_CELL[a.storage]
        out["roofline"]["achieved"] = round(BYTES_PER_CELL[a.storage] * work / (ms * 1e-3) / 1e9, 1)
        out["roofline"]["frac"] = round(out["roofline"]["achieved"] / HBM_PEAK_GBS, 4)
    if not a.no_sequential:
        tf = _sequential_fused(F, hip, insts)
        td = _sequential_dropin(F, insts)
        out["sequential_fused"] = {"seconds": round(sum(tf), 3), "gcell_per_s": round(work / sum(tf) / 1e9, 3)}
        out["sequential_dropin"] = {"seconds": round(sum(td), 3), "gcell_per_s": round(work / sum(td) / 1e9, 3)}
        # both sides end to end: each sequential solve includes its set-up, upload and free, the batch its
        # uploads and autotuning
        out["speedup_vs_sequential_dropin"] = round(sum(td) / e2e, 3)
        out["speedup_vs_sequential_fused"] = round(sum(tf) / e2e, 3)
        out["speedup_basis"] = "end to end on both sides (set-up and uploads included)"
    print(json.dumps(out), flush=True)


def bench_served(a):
    import threading

    import fdtd_amd as F
    from fdtd_amd import hip
    from fdtd_amd import workloads as W

    stream = W.served_stream(a.requests, f16_share=a.f16_share)
    reqs = [(t, _instance(F, w, st, int(a.load_zero_terms)), st) for t, w, st in stream]
    work = sum(x.params.nx * x.params.ny * x.params.nz * x.steps for _, x, _ in reqs)
    srv = F.Server(int(a.budget_gb * (1 << 30)), max_slots=a.slots, segment_steps=a.segment)
    ids = [None] * len(reqs)

    def client():
        t0 = time.perf_counter()
        for q, (t, x, _) in enumerate(reqs):
            d = t - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids[q] = srv.submit(x)
            if (q + 1) % 40 == 0:
                print(f"[served] submitted {q + 1}/{len(reqs)}", file=sys.stderr, flush=True)

    th = threading.Thread(target=client)
    th.start()
    th.join()
    while not srv.drain(30.0):
        st = srv.stats()
        print(f"[served] completed {st['completed']}/{len(reqs)}", file=sys.stderr, flush=True)
    st = srv.stats()
    finite = True
    for r in ids:
        s, info = srv.result(r)
        finite &= bool(np.isfinite(s).all()) and info["state"] == F.REQ_DONE
    srv.close()
    out = {
        "metric": "served solves/s (continuous batching, open-loop Poisson 40 req/s) with p50/p99 latency",
        "value": round(st["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1, "steps": a.requests, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 Debye sums) or f16 storage per request", "data": "synthetic (seeded phantoms)",
        "config": {"workload": "served_stream", "requests": len(reqs), "rate_per_s": 40.0,
                   "f16_requests": sum(st_ == "f16" for *_, st_ in reqs), "cell_updates": work,
                   "budget_gb": a.budget_gb, "slots": a.slots, "segment_steps": a.segment,
                   "parallelism": "single GPU"},
        "p50_latency_s": round(st["p50_latency_s"], 3), "p99_latency_s": round(st["p99_latency_s"], 3),
        "mean_latency_s": round(st["mean_latency_s"], 3),
        "gcell_per_s": round(st["cell_updates"] / max(st["t_last_done"] - st["t_first_submit"], 1e-9) / 1e9, 3),
        "peak_running": st["peak_running"], "peak_gb": round(st["peak_bytes"] / (1 << 30), 2),
        "completed": st["completed"], "failed": st["failed"], "results_finite": finite,
    }
    if not a.no_sequential:
        arrivals = [t for t, _, _ in reqs]

        def fifo(service):
            fin, lat, t = [], [], 0.0
            for arr, s in zip(arrivals, service):
                t = max(t, arr) + s
                fin.append(t)
                lat.append(t - arr)
            return {"seconds": round(fin[-1] - arrivals[0], 3), "solves_per_s": round(len(fin) / (fin[-1] - arrivals[0]), 3),
                    "p50_latency_s": round(_pct(lat, 0.5), 3), "p99_latency_s": round(_pct(lat, 0.99), 3)}

        td = _sequential_dropin(F, [x for _, x, _ in reqs])
        out["sequential_dropin"] = fifo(td)
        out["speedup_vs_sequential_dropin"] = round(out["value"] / out["sequential_dropin"]["solves_per_s"], 3)
        tf = _sequential_fused(F, hip, [x for _, x, _ in reqs])
        out["sequential_fused"] = fifo(tf)
        out["speedup_vs_sequential_fused"] = round(out["value"] / out["sequential_fused"]["solves_per_s"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
