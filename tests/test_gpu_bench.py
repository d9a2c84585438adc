"""bench.py on the GPU at small sizes: the JSON line keeps its contract
(metric, value, roofline with the dominant kernel's timing, config), and the
slab path runs a box phantom as z-slabs beside the unsplit grid."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    print("bench.py", *args, flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_bench_line_contract():
    d = _bench("--workload", "plain_small", "--steps", "32", "--warmup", "2", "--advance", "40",
               "--no-cpu-baseline", "--timing-launches", "8")
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 32 and d["value"] > 0 and d["probes_finite_nonzero"]
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and r["launches_timed"] == 64   # at least 64
    assert d["config"]["timed_steps"] == [42, 74] and r["timed_launch_steps"] == [74, 138]
    assert 0 < r["frac"] < 1 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3
    assert r["traffic"] is None      # no counter profile for this workload
    c = r["stream_ceiling"]
    assert c["gb_per_s"] > 0 and abs(c["frac"] - r["achieved"] / c["gb_per_s"]) < 1e-3
    assert d["config"]["workload"] == "plain_small" and d["config"]["grid"] == [64, 64, 64]


def test_bench_16bit_line_has_its_ceiling():
    d = _bench("--workload", "plain_small", "--storage", "f16", "--steps", "16", "--warmup", "2",
               "--no-cpu-baseline")
    r = d["roofline"]
    assert r["bytes_per_cell"] == 49 and d["dtype"].startswith("f16")
    c = r["stream_ceiling"]
    assert c["gb_per_s"] > 0 and abs(c["frac"] - r["achieved"] / c["gb_per_s"]) < 1e-3
    assert "131072 32-bit words" in c["what"]      # 64^3 cells, two per word


def test_bench_slabs_box_phantom():
    d = _bench("--workload", "slabs", "--grid", "96x64x60", "--slabs", "3", "--steps", "12", "--warmup", "2")
    assert d["config"]["workload"] == "box_phantom_slabs" and d["config"]["slab_planes"] == [20, 20, 20]
    assert d["value"] > 0 and d["unsplit_value"] > 0 and d["exchange_us_per_step"] > 0
    assert d["message_bytes_per_interface"]["up"] == 5 * 96 * 64 * 4   # five pitched planes (96 is a line multiple)


def test_stream_probe():
    import fdtd_amd as F
    print("stream probe, 1 M cells", flush=True)
    gbs, ms = F.stream_probe(1 << 20, 0, 4)
    assert 100.0 < gbs < 9000.0 and ms > 0.0
    with pytest.raises(F.FdtdError, match="-1"):
        F.stream_probe(0)
