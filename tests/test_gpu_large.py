"""The 320x320x880 grid at full size (50 tissue ids, a 10-cell Jz dipole 4
cells outside the body): 90 fused steps on the GPU equal the FP32
transcription of the reference bit for bit, in every field of the newest
level and in every probe sample.  By step 90 the dipole's wave has entered
the Debye body and reached the x = 0 Mur face.  The transcription runs on the
host's cores (about 10-20 s at 16 threads)."""
import time

import numpy as np
import pytest

import fdtd_amd as F
from fdtd_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 90


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
