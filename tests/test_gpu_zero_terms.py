"""Zero-term rows: the exact fp32 kernels skip loading E^{n-1} and D^{n-1}
on the half rows of a wave (32 cells) whose cells all have tau = 0 (their
Debye update multiplies those operands by +0).  Results must not change by
a single bit: every field of every ring level and every probe sample equals
the run that loads them (load_zero_terms = 1), in a context and in a batch,
also where the fields overflow (the reference's unstable Debye update,
fdtd_calculations.cu:144, doubles a vacuum field every step until it is
infinite and then NaN) and where fields hold -0 (the rows then load the
operands after all)."""
import numpy as np
import pytest

import fdtd_amd as F
from fdtd_amd import workloads as W
from tests.helpers import first_diff

pytestmark = pytest.mark.gpu


def _run(w, steps, load, debye=F.DEBYE_FIXED, shape_kw=None):
    ctx = F.Context(w.params, debye=debye, load_zero_terms=load, **(shape_kw or {}))
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, steps)
    ctx.run(steps)
    out = {(f, lv): ctx.get(f, lv) for f in range(9) for lv in ([0] if 3 <= f < 6 else [0, 1, 2])}
    series = ctx.probes()
    ctx.close()
    return out, series


def _same(a, b, tag):
    for key in a:
        x, y = a[key], b[key]
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{tag} {F.FIELDS[key[0]]} level {key[1]}: " \
                                                                       f"{first_diff(x, y)}"


@pytest.mark.parametrize("steps,shape_kw", [(300, {}), (50, {}), (300, {"tile_x": 128, "tile_y": 2, "kchunk": 6}),
                                            (50, {"tile_y": 8, "kchunk": 16})])
def test_zero_term_rows_change_no_bit(steps, shape_kw):
    w = W.plain_small(steps=steps)    # vacuum with a Debye sphere: most rows have tau = 0
    print(f"\nzero-term rows, plain_small {steps} steps, {shape_kw}", flush=True)
    a, sa = _run(w, steps, 0, shape_kw=shape_kw)
    b, sb = _run(w, steps, 1, shape_kw=shape_kw)
    _same(a, b, "fixed")
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64))
    if steps == 50:
        # the wave's far precursor holds negative E^n so small that d*E^n underflows to -0 (d ~ 7e-23),
        # the case in which a row loads the real operands after all
        e = np.concatenate([a[(f, 0)].ravel() for f in (6, 7, 8)])
        assert np.any((e < 0) & (e > -1e-24))


@pytest.mark.parametrize("shape_kw", [{}, {"tile_x": 128, "tile_y": 2, "kchunk": 6}])
def test_zero_term_half_rows_at_column_edges(shape_kw):
    """Vacuum with single dispersive columns on either side of and across the
    32-cell halves of a wave row (x = 31, 32, 33, 63, 64, 95, 96) and at the
    first and last columns of a partial last tile (x = 128, 149 of 150): each
    half wave decides its skip from its own 32 cells, and every bit equals
    the kernel that loads every row."""
    shape = (150, 21, 26)
    steps = 110
    params, jz = W.derive(shape, steps, 2.45e9, 2.0e-3, 0.99)
    ids = np.zeros(shape[::-1], np.uint8)
    for band, x in enumerate((31, 32, 33, 63, 64, 95, 96)):
        ids[:, 3 * band:3 * band + 3, x] = 1
    ids[:13, 18:, 128] = 1
    ids[13:, 18:, 149] = 1
    table = np.asarray([(0.0, 1.0, 1.0, 0.0), (0.8, 55.0, 40.0, 8e-12)], np.float32)
    src = np.asarray([[48, 10, 13], [112, 10, 13]], np.int32)
    probes = np.asarray([[8, 32, 10, 13], [8, 31, 4, 10], [8, 128, 19, 6], [8, 149, 19, 20]], np.int32)
    w = W.Workload("column_edges", shape, params, ids, table, src, jz[:steps].copy(), probes, steps)
    print(f"\nzero-term half rows at column edges, {shape} {steps} steps, {shape_kw}", flush=True)
    a, sa = _run(w, steps, 0, shape_kw=shape_kw)
    b, sb = _run(w, steps, 1, shape_kw=shape_kw)
    _same(a, b, "columns")
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64))
    assert np.isfinite(sa).all() and (np.abs(sa).max(axis=0) > 0).all()   # the wave reached every probe


def test_zero_term_rows_through_overflow():
    """The reference's Debye mode: vacuum fields grow 2x per step, overflow to
    inf near step 130 and turn NaN; the skipped rows must fall back to the
    real operands and give the same bits (NaN payloads included)."""
    w = W.plain_small(steps=220)
    print("\nzero-term rows through overflow, reference Debye mode, 220 steps", flush=True)
    a, sa = _run(w, 220, 0, debye=F.DEBYE_REF)
    b, sb = _run(w, 220, 1, debye=F.DEBYE_REF)
    assert not np.isfinite(a[(8, 0)]).all()
    _same(a, b, "ref")
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64))


def test_zero_term_rows_in_a_batch():
    ws = W.batched_set(count=6, seed=777)
    steps = 60
    series = []
    for load in (0, 1):
        insts = [F.Instance(w.params, F.make_ext(load_zero_terms=load), w.ids, w.table, w.sources, w.jz,
                            min(w.steps, steps), w.probes) for w in ws]
        with F.Batch(int(sum(x.nbytes() for x in insts) * 1.05) + (32 << 20)) as b:
            ids = [b.add(x) for x in insts]
            b.run(steps)
            series.append([(b.probes(i), b.get(i, 8), b.get(i, 2, 1)) for i in ids])
    for q, (x, y) in enumerate(zip(*series)):
        for u, v in zip(x, y):
            assert np.array_equal(np.ascontiguousarray(u).view(np.uint8), np.ascontiguousarray(v).view(np.uint8)), q
