#!/usr/bin/env python3
"""Throughput of the fused FDTD time step on one MI355X.

A "step" is one full FDTD time step of the grid: the fused H/D/source/Debye-E
kernel, the Mur face kernel, the probe kernel and (every 100 steps) the
z-slice capture.  The default workload is the 256^3 eight-ellipsoid Debye
phantom with a Jz sheet source at 900 MHz, 3 probes and a z-slice every 100
steps (plain_256 in fdtd_amd/workloads.py), 2000 steps.

Prints one JSON line: Gcell-updates/s over the timed steps, the fused
kernel's roofline (97 algorithmic bytes per cell-update over its mean
launch duration, timed with HIP events on its stream) and the CPU baseline
(the oracle's FP32 transcription, OpenMP on this host's cores, on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CELL = {"f32": 97, "f16": 49, "bf16": 49}  # 24 state words + 1 material byte per cell-update
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plain_256", choices=["plain_256", "large_phantom", "plain_small"])
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"],
                    help="field storage (f32 reproduces the reference's arithmetic bit for bit)")
    ap.add_argument("--cpu-steps", type=int, default=4, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run one device-to-device copy of this many bytes first (HBM counter calibration)")
    ap.add_argument("--traffic-kib", type=float, nargs=2, default=None,
                    help="measured FETCH_SIZE and WRITE_SIZE per fused launch (KiB, already scaled)")
    return ap.parse_args()


def build_workload(name, steps):
    from fdtd_amd import workloads as W
    if name == "plain_256":
        return W.plain_256(steps=steps)
    if name == "large_phantom":
        return W.large_phantom(steps=steps)
    return W.plain_small(steps=steps)


def cpu_baseline(w, steps):
    """The oracle's FP32 transcription on this host's cores (a baseline only)."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    sim = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0, prec=32)
    sim.set_materials(*w.material_arrays())
    sim.set_sources(w.sources, w.jz, jz_mode=1)
    sim.step(1)  # first touch of the arrays
    t0 = time.perf_counter()
    sim.step(steps)
    dt = time.perf_counter() - t0
    nx, ny, nz = w.shape
    return {"value": round(nx * ny * nz * steps / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{w.name} {nx}x{ny}x{nz}, {steps} steps of the oracle FP32 transcription "
                      f"(OpenMP over z, {threads} threads), {dt:.2f} s"}


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

    total = a.warmup + a.steps
    w = build_workload(a.workload, total)
    nx, ny, nz = w.shape
    cells = nx * ny * nz

    stream = hip.Stream()
    storage = {"f32": F.STORE_F32, "f16": F.STORE_F16, "bf16": F.STORE_BF16}[a.storage]
    bpc = BYTES_PER_CELL[a.storage]
    ctx = F.Context(w.params, storage=storage)
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, total)
    if w.slice_every:
        ctx.set_slices(w.slice_k, w.slice_every, total // w.slice_every + 1)
    ctx.run(a.warmup, stream.handle)
    stream.sync()
    ctx.set_timing(a.steps)

    e0, e1 = hip.Event(), hip.Event()
    hip.sync()
    e0.record(stream)
    ctx.run(a.steps, stream.handle)
    e1.record(stream)
    e1.sync()
    hip.sync()
    ms = e0.elapsed_ms(e1)
    kms, launches = ctx.kernel_time()
    probes = ctx.probes()
    ok = bool(abs(probes).max() > 0) and bool((probes == probes).all())

    value = cells * a.steps / (ms * 1e-3) / 1e9
    kavg_s = kms / launches * 1e-3
    achieved = bpc * cells / kavg_s / 1e9
    traffic = None
    if a.traffic_kib:
        traffic = int((a.traffic_kib[0] + a.traffic_kib[1]) * 1024)
    shape = ctx.launch_shape()
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
        "data": "synthetic (seeded phantom)",
        "config": {"workload": w.name, "grid": [nx, ny, nz], "cells": cells, "materials": int(w.table.shape[0]),
                   "sources": int(w.sources.shape[0]), "probes": int(w.probes.shape[0]),
                   "slice_every": w.slice_every, "launch": shape, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "fused_step", "kernel_avg_ms": round(kms / launches, 5),
                     "bytes_per_cell": bpc},
        "probes_finite_nonzero": ok,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_steps)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
