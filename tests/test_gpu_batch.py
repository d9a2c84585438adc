"""Batches of uneven grids: one persistent fused launch per storage form and
step over all running grids, replayed from captured HIP graphs.

A grid's results in a batch must be those of a single context with the same
inputs, bit for bit, in every storage form: the per-cell arithmetic does not
depend on the batch, the tile shape or the step at which the grid joined.
The single context is itself bit-exact against the oracle's FP32
transcription (test_gpu_parity.py), so this pins the batch to the oracle."""
import numpy as np
import pytest

import fdtd_amd as F
from tests.helpers import EPS0, MU0, bits_equal, first_diff, random_materials, random_rp, random_sources

# fp16 grids here run a few dozen steps with probes far from their sources, whose precursor
# peaks below fp16's normal range: the precision warning is expected and not what these
# bit-for-bit comparisons test
pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore::fdtd_amd.PrecisionWarning")]

DX = 2e-3
DT = 0.99 * DX / (3.0e8 * np.sqrt(3.0))


def waveform(rng, n=600):
    t = np.arange(n, dtype=np.float64)
    t0, w = rng.uniform(20, 40), rng.uniform(5, 12)
    return (-(t - t0) / w * np.exp(-((t - t0) / w) ** 2) * rng.uniform(0.5, 3.0)).astype(np.float32)


def random_probes(rng, shape, count):
    """Random fields and cells; every other probe is an E component on a
    boundary face, where the Mur faces write in the probes' own launch."""
    nx, ny, nz = shape
    p = np.stack([rng.integers(0, 9, count), rng.integers(0, nx, count), rng.integers(0, ny, count),
                  rng.integers(0, nz, count)], 1).astype(np.int32)
    for r in range(1, count, 2):
        p[r, 0] = 6 + rng.integers(0, 3)
        ax = int(rng.integers(0, 3))
        p[r, 1 + ax] = 0 if rng.integers(0, 2) == 0 else shape[ax] - 1
    return p


def make_case(seed, shape, nmat, steps, storage=F.STORE_F32, jz_mode=F.JZ_STEP, rp=False, nprobe=6,
              debye=F.DEBYE_FIXED, jz_limit=0):
    rng = np.random.default_rng(seed)
    ids, table = random_materials(rng, shape, nmat)
    src = random_sources(rng, shape, 2)
    jz = waveform(rng)
    probes = random_probes(rng, shape, nprobe)
    rps = random_rp(rng, shape) if rp else None
    params = F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0)
    ext = F.make_ext(debye=debye, jz_mode=jz_mode, jz_limit=jz_limit, storage=storage)
    return F.Instance(params, ext, ids, table, src, jz, steps, probes, rps)


def single(inst):
    """The same grid alone in a fused context."""
    e = inst.ext
    ctx = F.Context(inst.params, debye=e.debye_mode, jz_mode=e.jz_index_mode, jz_limit=e.jz_limit, storage=e.storage)
    ctx.set_materials(inst.ids, inst.table)
    if inst.rp is not None:
        ctx.set_rp(inst.rp)
    ctx.set_sources(inst.src, inst.jz)
    if inst.nprobe:
        ctx.set_probes(inst.probes, inst.steps)
    ctx.run(inst.steps)
    fields = {(f, lv): ctx.get(f, lv) for f in range(9) for lv in ([0] if 3 <= f < 6 else [0, 1, 2])}
    return fields, ctx.probes()


def assert_same(batch, i, inst, ref, label):
    fields, series = ref
    done, total = batch.status(i)
    assert (done, total) == (inst.steps, inst.steps), label
    for (f, lv), want in fields.items():
        got = batch.get(i, f, lv)
        assert bits_equal(got, want), f"{label}: {F.FIELDS[f]} level {lv}: {first_diff(got, want)}"
    got = batch.probes(i)
    assert got.shape == series.shape, label
    assert np.array_equal(got.view(np.uint64), series.view(np.uint64)), f"{label}: probe series differ"


CASES = [  # shape, nmat, steps, storage, jz mode, rp
    ((20, 24, 28), 4, 17, F.STORE_F32, F.JZ_STEP, False),
    ((33, 17, 40), 12, 30, F.STORE_F32, F.JZ_HOLD3, True),
    ((70, 9, 12), 6, 25, F.STORE_F16, F.JZ_STEP, False),
    ((16, 16, 16), 5, 9, F.STORE_F32_FAST, F.JZ_STEP, False),
    ((48, 40, 20), 8, 40, F.STORE_BF16, F.JZ_STEP, True),
    ((65, 5, 7), 4, 33, F.STORE_F32, F.JZ_STEP, False),
]


def test_batch_equals_single_grids():
    insts = [make_case(100 + q, *c) for q, c in enumerate(CASES)]
    refs = [single(x) for x in insts]
    b = F.Batch(1 << 30)
    b.set_shape(tile_y=4, kchunk=8, graph_steps=7)
    ids = [b.add(x) for x in insts[:4]]
    print(f"\n[batch] 4 grids added, run 11 steps", flush=True)
    b.run(11)
    ids += [b.add(x) for x in insts[4:]]   # these start at batch step 11
    print("[batch] 2 grids added at step 11, run 50 steps", flush=True)
    b.run(50)
    info = b.info()
    assert info["instances"] == 6 and info["running"] == 0 and info["step"] == 61
    for q, (i, x, r) in enumerate(zip(ids, insts, refs)):
        assert_same(b, i, x, r, f"case {q} {CASES[q][0]}")
    b.close()


def test_batch_slot_and_memory_reuse():
    a = make_case(7, (24, 20, 18), 5, 12)
    c = make_case(8, (40, 12, 22), 7, 15, storage=F.STORE_F16)
    d = make_case(9, (18, 30, 26), 6, 21, jz_mode=F.JZ_HOLD3)
    b = F.Batch(256 << 20)
    b.set_shape(tile_y=2, kchunk=4, graph_steps=5)
    ia, ic = b.add(a), b.add(c)
    print("\n[batch] reuse: run 20", flush=True)
    b.run(20)
    assert_same(b, ia, a, single(a), "first grid")
    used = b.info()["used_bytes"]
    b.remove(ia)
    assert b.info()["used_bytes"] < used
    idd = b.add(d)           # takes the freed slot and region
    assert idd == ia
    b.run(25)
    assert_same(b, ic, c, single(c), "grid beside the reused slot")
    assert_same(b, idd, d, single(d), "grid in the reused slot")
    b.close()


def test_batch_autotune_restores_state():
    insts = [make_case(200 + q, s, 5, n) for q, (s, n) in enumerate([((30, 30, 30), 14), ((50, 20, 36), 11),
                                                                      ((12, 40, 24), 16)])]
    refs = [single(x) for x in insts]
    b = F.Batch(512 << 20)
    ids = [b.add(x) for x in insts]
    print("\n[batch] autotune", flush=True)
    ty, kc = b.autotune(trial_steps=3)
    assert ty in (2, 4, 8) and kc in (3, 4, 6, 8, 16)
    info = b.info()
    assert (info["tile_y"], info["kchunk"], info["step"]) == (ty, kc, 0)
    b.run(16)
    for q, (i, x, r) in enumerate(zip(ids, insts, refs)):
        assert_same(b, i, x, r, f"tuned grid {q}")
    b.close()


def test_batch_rejects_what_does_not_fit():
    b = F.Batch(1 << 20)
    x = make_case(3, (64, 64, 64), 4, 10)
    with pytest.raises(F.FdtdError, match="-5"):
        b.add(x)
    b.close()


def test_dropin_solve_equals_fused():
    """The drop-in entry points driven like the reference's main loop give the
    fused context's probe series bit for bit (jz within the 512-value limit)."""
    for q, jm in enumerate([F.JZ_HOLD3, F.JZ_STEP]):
        inst = make_case(300 + q, (26, 19, 31), 6, 23, jz_mode=jm, rp=True, nprobe=5, jz_limit=512)
        got = F.dropin_solve(inst)
        want = single(inst)[1]
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), f"jz mode {jm}"
    # rejected before any device work: a waveform length without its values, a material id
    # past the table
    bad = make_case(310, (20, 18, 16), 4, 5)
    bad.c.jz = None
    with pytest.raises(F.FdtdError, match="-1"):
        F.dropin_solve(bad)
    bad = make_case(311, (20, 18, 16), 4, 5)
    bad.ids[3] = 4
    with pytest.raises(F.FdtdError, match="-2"):
        F.dropin_solve(bad)


def test_batch_reference_debye_mode():
    """Grids in the reference's Debye mode (d = 2b - sigma dt^2, unstable but
    short runs stay finite) and the stable default in one batch: each equals
    its single context bit for bit, fp32 and fp16."""
    insts = [make_case(400, (22, 18, 26), 6, 24, debye=F.DEBYE_REF),
             make_case(401, (30, 14, 20), 5, 19, storage=F.STORE_F16, debye=F.DEBYE_REF),
             make_case(402, (18, 26, 16), 4, 21)]
    refs = [single(x) for x in insts]
    b = F.Batch(1 << 28)
    print("\n[batch] REF and FIXED Debye modes together", flush=True)
    ids = [b.add(x) for x in insts]
    b.run(30)
    for q, (i, x, r) in enumerate(zip(ids, insts, refs)):
        assert_same(b, i, x, r, f"debye case {q}")
    assert np.isfinite(refs[0][1]).all()
    b.close()
