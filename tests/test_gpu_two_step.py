"""Two time steps per fused launch (FdtdExt.time_steps = 2) against one step
per launch.  The two-step kernel recomputes step n+1 on tiles grown by one
cell, applies that step's six Mur faces inside the tile and writes level n+2
into a spare slot; every field of every ring level and every probe sample
must still agree bit for bit.  Shapes cover widths across one and several
62-cell tile columns (x faces at either end of a tile), rows across tile
boundaries, z-chunks of 1 to 16 planes, sources on the upper faces, odd step
counts (a one-step launch at the end), and energy sums and slices that force
single steps in between."""
import numpy as np
import pytest

import fdtd_amd as F
from fdtd_amd import workloads as W
from oracle import oracle as O
from tests.helpers import EPS0, MU0, first_diff, random_materials, random_rp, random_sources

pytestmark = pytest.mark.gpu

DT = 3.8105118e-12
DX = 2.0e-3


def _jz(n=60):
    t = np.arange(n, dtype=np.float64)
    return (-(t - 12) / 4 * np.exp(-((t - 12) / 4) ** 2) * 2.0).astype(np.float32)


def _sources(rng, shape):
    nx, ny, nz = shape
    face = np.asarray([[nx - 1, ny // 2, nz // 2], [nx // 2, ny - 1, 0], [2, 1, nz - 1]], np.int32)
    return np.concatenate([random_sources(rng, shape, 3), face])


def _probes(shape):
    nx, ny, nz = shape
    return np.asarray([[8, nx // 2, ny // 2, nz // 2], [7, 0, ny // 2, nz // 2], [6, nx - 1, ny // 3, 1],
                       [2, 1, 1, 0], [7, nx - 2, ny - 2, nz - 1], [8, nx // 3, 0, nz // 2], [6, 1, ny - 1, nz // 3],
                       [1, nx - 1, ny - 1, nz - 1], [0, nx // 2, 1, 1]], np.int32)


def run(shape, time_steps, steps, seed, kchunk=0, energy=0, slices=0, xcd_remap=None, probes=None):
    rng = np.random.default_rng(seed)
    ids, table = random_materials(rng, shape, 6)
    ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0), kchunk=kchunk, time_steps=time_steps,
                    xcd_remap=xcd_remap)
    ctx.set_materials(ids, table)
    ctx.set_rp(random_rp(rng, shape))
    ctx.set_sources(_sources(rng, shape), _jz())
    ctx.set_probes(_probes(shape) if probes is None else probes, steps)
    if energy:
        ctx.set_energy(energy, steps)
    if slices:
        ctx.set_slices(shape[2] // 2, slices, steps)
    ctx.set_timing(steps)
    ctx.run(steps)
    fields = {(f, lv): ctx.get(f, lv) for f in range(9) for lv in ([0] if 3 <= f < 6 else [0, 1, 2])}
    out = dict(fields=fields, probes=ctx.probes(), n=ctx.n, timed_steps=ctx.timed_steps())
    if energy:
        out["energy"] = ctx.energy()
    if slices:
        out["slices"] = ctx.slices()
    ctx.close()
    return out


def _same(a, b):
    assert a["n"] == b["n"]
    for key in a["fields"]:
        x, y = a["fields"][key], b["fields"][key]
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{F.FIELDS[key[0]]} level {key[1]}: {first_diff(x, y)}"
    assert np.array_equal(a["probes"].view(np.uint64), b["probes"].view(np.uint64)), \
        np.argwhere(a["probes"].view(np.uint64) != b["probes"].view(np.uint64))[:5]


@pytest.mark.parametrize("shape,kchunk,steps", [
    ((37, 29, 23), 0, 25), ((130, 17, 9), 4, 24), ((63, 7, 8), 3, 21), ((64, 33, 40), 16, 30), ((5, 6, 7), 0, 17),
    ((200, 20, 30), 5, 26), ((125, 9, 11), 1, 14), ((62, 12, 10), 2, 20), ((124, 6, 6), 0, 13), ((3, 3, 3), 0, 9)])
def test_two_steps_equal_single_steps(shape, kchunk, steps):
    print(f"\n[two-step] {shape} kc={kchunk} steps={steps}", flush=True)
    a = run(shape, 2, steps, 5, kchunk)
    b = run(shape, 1, steps, 5, kchunk)
    assert a["timed_steps"] == 2 and b["timed_steps"] == 1   # the first launch ran two steps
    _same(a, b)
    assert np.isfinite(a["probes"]).all() and np.abs(a["probes"]).max() > 0


def test_two_steps_with_energy_and_slices():
    """Energy sums every 3 steps and slices every 4 put single steps between
    the pairs; their series must be unchanged too."""
    shape = (70, 26, 19)
    print(f"\n[two-step] {shape} energy/slices", flush=True)
    a = run(shape, 2, 31, 8, energy=3, slices=4)
    b = run(shape, 1, 31, 8, energy=3, slices=4)
    assert a["timed_steps"] == 2
    _same(a, b)
    assert np.array_equal(a["energy"][0], b["energy"][0])
    assert np.array_equal(a["energy"][1].view(np.uint64), b["energy"][1].view(np.uint64))
    sa, sb = a["slices"], b["slices"]
    assert list(sa[0]) == list(sb[0]) == list(range(4, 32, 4))
    for x, y in zip(sa[1], sb[1]):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))


def test_two_steps_h_probes_fall_back():
    """H probes need H^{n+1/2}, which two-step launches never store: such a
    context steps one at a time, with the same results."""
    shape = (40, 21, 19)
    pr = np.asarray([[3, 20, 10, 9], [5, 1, 1, 1], [8, 10, 10, 10]], np.int32)
    a = run(shape, 2, 12, 3, probes=pr)
    b = run(shape, 1, 12, 3, probes=pr)
    assert a["timed_steps"] == 1
    _same(a, b)


@pytest.mark.parametrize("xcd_remap", [0, 1, 2, 3, 4])
def test_block_orders(xcd_remap):
    """Every block -> tile order, with one and two steps per launch, against
    the identity order (a permutation of tiles changes no result).  3 x 9
    tiles per layer in 10 layers: every order has full groups of 8 units
    and a partial last group."""
    shape = (130, 33, 40)
    ref = run(shape, 1, 10, 4, kchunk=4, xcd_remap=0)
    for ts in (1, 2):
        _same(run(shape, ts, 10, 4, kchunk=4, xcd_remap=xcd_remap), ref)


def test_two_steps_against_the_transcription():
    """A two-step context against the oracle's FP32 transcription of the
    reference directly (Debye tissues, random Mur planes, face sources)."""
    shape = (70, 30, 25)
    steps = 40
    rng = np.random.default_rng(21)
    ids, table = random_materials(rng, shape, 6)
    rps = random_rp(rng, shape)
    src = _sources(rng, shape)
    probes = _probes(shape)
    ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0), time_steps=2)
    ctx.set_materials(ids, table)
    ctx.set_rp(rps)
    ctx.set_sources(src, _jz())
    ctx.set_probes(probes, steps)
    ctx.set_timing(1)
    ctx.run(steps)
    assert ctx.timed_steps() == 2
    sim = O.Sim(shape, DT, DX, DX, DX, MU0, EPS0, prec=32)
    t = table[ids.astype(np.int64)]
    sim.set_materials(t[..., 0], t[..., 1], t[..., 2], t[..., 3])
    sim.set_rp(rps)
    sim.set_sources(src, _jz(), jz_mode=1)
    ref = sim.run(steps, probes)
    assert np.array_equal(ctx.probes().view(np.uint64), ref.view(np.uint64))
    for f in range(9):
        x, y = ctx.get(f), sim.get(f)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{F.FIELDS[f]}: {first_diff(x, y)}"
    ctx.close()


def test_two_steps_plain_256_full_horizon():
    """The 256^3 Debye phantom over its full 2000 steps (3 Ez probes, a
    z-slice every 100 steps, which puts one single step in every 100): two
    steps per launch against one, every field, level, probe and slice."""
    steps = 2000
    w = W.plain_256(steps=steps)

    def go(ts):
        ctx = F.Context(w.params, time_steps=ts)
        ctx.set_materials(w.ids, w.table)
        ctx.set_sources(w.sources, w.jz)
        ctx.set_probes(w.probes, steps)
        ctx.set_slices(w.slice_k, w.slice_every, steps // w.slice_every)
        ctx.set_timing(1)
        ctx.run(steps)
        out = dict(fields={(f, lv): ctx.get(f, lv) for f in range(9) for lv in ([0] if 3 <= f < 6 else [0, 1, 2])},
                   probes=ctx.probes(), n=ctx.n, timed_steps=ctx.timed_steps(), slices=ctx.slices())
        ctx.close()
        return out

    print("\nplain_256, 2000 steps, two-step against one-step", flush=True)
    a, b = go(2), go(1)
    assert a["timed_steps"] == 2
    _same(a, b)
    assert np.abs(a["probes"]).max() > 0 and np.isfinite(a["probes"]).all()
    assert list(a["slices"][0]) == list(b["slices"][0]) == list(range(100, steps + 1, 100))
    assert np.array_equal(a["slices"][1].view(np.uint32), b["slices"][1].view(np.uint32))
