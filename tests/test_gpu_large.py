"""The two large benchmark grids at full size against the FP32 transcription
of the reference, bit for bit, in every field of the newest level and in
every probe sample:
  * 320x320x880 (50 tissue ids, a 10-cell Jz dipole 4 cells outside the
    body), 300 steps: by then the dipole's wave has crossed into the Debye
    body and reached the x = 0 Mur face (about 45 s of host time on 16
    threads);
  * 256^3 (8 nested Debye ellipsoids, a 63504-cell Jz sheet), its full
    2000 steps with the 3 probes and the z-slice every 100 steps (about 60 s
    of host time).
Both run the fused kernel with zero-term rows, as the bench does."""
import time

import numpy as np
import pytest

import fdtd_amd as F
from fdtd_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 300


def test_large_phantom_bit_equal_to_the_transcription():
    w = W.large_phantom(steps=STEPS)
    print(f"\nlarge_phantom {w.shape}, {STEPS} steps", flush=True)
    ctx = F.Context(w.params)
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, STEPS)
    ctx.run(STEPS)
    got = ctx.probes()
    t0 = time.perf_counter()
    sim = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0, prec=32)
    sim.set_materials(*w.material_arrays())
    sim.set_sources(w.sources, w.jz, jz_mode=1)
    ref = sim.run(STEPS, w.probes)
    print(f"transcription: {time.perf_counter() - t0:.1f} s", flush=True)
    assert np.array_equal(got, ref)
    assert np.abs(ref[:, 1]).max() > 0   # the probe beside the dipole saw the pulse
    for f in range(9):
        a = ctx.get(f).reshape(-1)
        b = sim.get(f).reshape(-1)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), F.FIELDS[f]
    # the wave reached the Debye body (id > 0) next to the dipole
    ez = ctx.get(8)
    assert np.abs(ez[w.ids > 0]).max() > 0
    ctx.close()


def test_plain_256_full_horizon_bit_equal_to_the_transcription():
    steps = 2000
    w = W.plain_256(steps=steps)
    print(f"\nplain_256 {w.shape}, {steps} steps", flush=True)
    ctx = F.Context(w.params)
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, steps)
    ctx.set_slices(w.slice_k, w.slice_every, steps // w.slice_every)
    ctx.run(steps)
    got = ctx.probes()
    t0 = time.perf_counter()
    sim = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0, prec=32)
    sim.set_materials(*w.material_arrays())
    sim.set_sources(w.sources, w.jz, jz_mode=1)
    parts = []
    for done in range(0, steps, 250):
        parts.append(sim.run(250, w.probes))
        print(f"transcription: step {done + 250}, {time.perf_counter() - t0:.1f} s", flush=True)
    ref = np.concatenate(parts)
    assert np.array_equal(got, ref)
    assert np.abs(ref).max() > 0 and np.isfinite(ref).all()
    for f in range(9):
        a = ctx.get(f).reshape(-1)
        b = sim.get(f).reshape(-1)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), F.FIELDS[f]
    slice_steps, data = ctx.slices()
    assert list(slice_steps) == list(range(100, steps + 1, 100))
    for f in range(9):
        assert np.array_equal(data[-1][f].view(np.uint32), ctx.get(f)[w.slice_k].view(np.uint32)), F.FIELDS[f]
    ctx.close()
