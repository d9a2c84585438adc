"""bench.py's contract on the CPU: more than one GPU is refused with a
message, the committed counter profile yields the traffic figure of the
default launch shape, and the CPU baseline carries its fields."""
import json
import os
import subprocess
import sys

import bench
from fdtd_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_batched_fixed_shape_takes_64_wide_tiles():
    """--shape for the batch fixes tile rows and planes of its 64-wide tiles;
    any other width is refused before the device is touched."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "batched", "--shape", "128x2x6"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "64xTYxKC" in r.stderr


def test_profiled_traffic_matches_the_launch_shape():
    """Each figure comes from the section of its own workload, launch shape,
    block order and kernel form (zero-term rows, or every row loading),
    wherever it sits."""
    alg = 97 * 320 * 320 * 880
    # the headline workload at the library's default launch shape and block order has its own counter passes
    tl, pl = bench.profiled_traffic("large_phantom", "f32", "64x4x16", order=3)
    assert tl and 1.0 < tl / alg < 1.1 and pl.endswith("fused_step_large_phantom.md")
    # the identity order re-reads halo columns on other XCDs
    ti, pi = bench.profiled_traffic("large_phantom", "f32", "64x4x16", order=0)
    assert ti and 1.2 < ti / alg < 1.4 and pi == pl
    # 4-plane chunks read more (a z-chunk prologue per 4 planes), the loading kernel more still
    t4, _ = bench.profiled_traffic("large_phantom", "f32", "64x4x4")
    t4l, _ = bench.profiled_traffic("large_phantom", "f32", "64x4x4", loading=True)
    assert t4 and t4l and ti < t4 < t4l < 1.8 * alg
    assert bench.profiled_traffic("large_phantom", "f32", "64x4x4", order=3) == (None, None)
    # the 256^3 counters of the autotuned shape were taken with the loading kernel only
    t, p = bench.profiled_traffic("plain_256", "f32", "128x2x6", loading=True)
    t0, _ = bench.profiled_traffic("plain_256", "f32", "64x4x4", loading=True)
    assert t and t0 and 1.3 < t / (97 * 256 ** 3) < 1.7 and t0 > t and p.endswith(".md")
    assert bench.profiled_traffic("plain_256", "f32", "128x2x6") == (None, None)
    # 256^3 at its default shape, with zero-term rows, in both block orders
    t8, p8 = bench.profiled_traffic("plain_256", "f32", "64x4x8", order=3)
    t8i, _ = bench.profiled_traffic("plain_256", "f32", "64x4x8", order=0)
    assert t8 and t8i and 1.0 < t8 / (97 * 256 ** 3) < t8i / (97 * 256 ** 3) < 1.7
    assert p8.endswith("fused_step_plain_256.md")
    # the RIKEN phantom at its default shape
    tr, pr = bench.profiled_traffic("riken", "f32", "64x4x16", order=3)
    assert tr and 1.0 < tr / (97 * 297 * 522 * 1719) < 1.2 and pr.endswith("fused_step_riken.md")
    # a shape or order without counters, or 16-bit storage, gives none
    assert bench.profiled_traffic("plain_256", "f32", "64x2x3") == (None, None)
    assert bench.profiled_traffic("plain_256", "f32", "64x4x8", order=2) == (None, None)
    assert bench.profiled_traffic("plain_256", "f16", "128x2x6", loading=True) == (None, None)
    # fp16 storage at 320x320x880 has its own passes (49 B/cell model); bf16 has none
    t16, p16 = bench.profiled_traffic("large_phantom", "f16", "64x4x16", order=3)
    assert t16 and 1.1 < t16 / (49 * 320 * 320 * 880) < 1.4 and p16.endswith("fused_step_large_phantom_f16.md")
    assert t16 != tl
    assert bench.profiled_traffic("large_phantom", "bf16", "64x4x16", order=3) == (None, None)


def test_headline_defaults():
    """With no flags the bench measures the 320x320x880 phantom at the
    library's default launch shape after an untimed advance, and times at
    least 64 fused launches for the roofline."""
    import argparse
    old = sys.argv
    try:
        sys.argv = ["bench.py"]
        a = bench.parse()
    finally:
        sys.argv = old
    assert isinstance(a, argparse.Namespace)
    assert a.workload == "large_phantom" and not a.autotune and a.advance >= 200 and a.timing_launches >= 64


def test_cpu_baseline_fields():
    w = W.plain_small(steps=8)
    b = bench.cpu_baseline(w, 4)
    assert b["value"] > 0 and b["unit"] == "Gcell-updates/s" and b["kind"] == "port"
    assert b["cores"] >= 1 and b["single_core"]["cores"] == 1 and b["single_core"]["value"] > 0
    assert "plain_small" in b["sample"] and b["host"]["nproc"] >= 1


def test_long_runs_script_imports():
    """tests/long_runs.py (run by hand on a GPU host) stays importable and
    exposes its three runs."""
    import tests.long_runs as L
    assert callable(L.small) and callable(L.served) and callable(L.large)
    assert L.rel_l2([1.0, 2.0], [1.0, 2.0]) == 0.0
