"""Sources on the upper faces and the top plane.  updateSources writes Dz
at any listed cell (fdtd_calculations.cu:198-211), also where the D update
does not reach: x = nx-1, y = ny-1, z = nz-1 (the Ez update then reads it;
Mur overwrites the x and y faces' Ez, but not the top plane's).  Every path
of this library accepts the same cells (1 <= x < nx, 1 <= y < ny,
0 <= z < nz) and gives the oracle's result bit for bit: the fused context in
several launch shapes (the x-high face folded into the fused tile or left to
the Mur kernel), the drop-in entry points, a batch, the 16-bit pair kernel
against the one-cell kernel, and the CLI from a parameter file."""
import os

import numpy as np
import pytest

import fdtd_amd as F
from oracle import oracle as O
from tests.helpers import oracle_sim, random_materials, random_rp, random_state
from tests.test_gpu_batch import single
from tests.test_gpu_parity import DT, DX, Legacy, _compare_all, _params

pytestmark = pytest.mark.gpu


def face_sources(shape):
    nx, ny, nz = shape
    return np.asarray([[nx - 1, ny // 2, nz // 2], [nx // 3, ny - 1, 2], [nx // 2, ny // 3, nz - 1],
                       [nx - 1, ny - 1, nz - 1], [1, 1, 0], [nx // 2, ny // 2, nz // 2]], np.int32)


def _jz(rng):
    return rng.uniform(-2, 2, 20).astype(np.float32)


@pytest.mark.parametrize("shape,tile_y,kchunk,tile_x", [((37, 29, 23), 4, 0, 64), ((65, 20, 30), 8, 0, 64),
                                                        ((100, 13, 11), 8, 0, 128), ((40, 21, 19), 4, 20, 64)])
def test_face_sources_fused_and_legacy_bit_exact(shape, tile_y, kchunk, tile_x):
    print(f"face sources {shape} ty={tile_y} kc={kchunk} tx={tile_x}", flush=True)
    rng = np.random.default_rng(sum(shape))
    ids, table = random_materials(rng, shape)
    rps = random_rp(rng, shape)
    src = face_sources(shape)
    jz = _jz(rng)
    st = random_state(rng, shape)
    sim = oracle_sim(shape, DT, DX)
    t = table[ids.astype(np.int64)]
    sim.set_materials(t[..., 0], t[..., 1], t[..., 2], t[..., 3])
    sim.set_rp(rps)
    sim.set_sources(src, jz, jz_mode=0)
    for (f, lv), a in st.items():
        sim.set(f, lv, a)
    with F.Context(_params(shape), jz_mode=F.JZ_HOLD3, tile_y=tile_y, kchunk=kchunk, tile_x=tile_x) as ctx:
        ctx.set_materials(ids, table)
        ctx.set_rp(rps)
        ctx.set_sources(src, jz)
        for (f, lv), a in st.items():
            ctx.set(f, lv, a)
        leg = Legacy(shape, table, ids, rps, src, jz, O.DEBYE_FIXED)
        leg.set_state(st)
        for n in range(7):
            ctx.step()
            leg.step((n // 3) * 3)
            sim.step()
        _compare_all(ctx.get, sim.get, f"fused {shape}")
        _compare_all(leg.get, sim.get, f"legacy {shape}")
    nx, ny, nz = shape
    # the top-plane source changed Ez there (no Mur face on Ez at z = nz-1)
    dz, ez = sim.get(2), sim.get(8)
    x, y = src[2, 0], src[2, 1]
    assert dz[nz - 1, y, x] != 0.0 and ez[nz - 1, y, x] != 0.0


@pytest.mark.parametrize("storage", [F.STORE_F16, F.STORE_BF16])
def test_face_sources_pairs_equal_single_cells(storage):
    """The x-high face source on either cell of a pair."""
    for shape in ((129, 12, 14), (130, 17, 9)):
        rng = np.random.default_rng(5)
        ids, table = random_materials(rng, shape, 6)
        src = face_sources(shape)
        jz = _jz(rng)
        out = []
        for lc in (2, 1):
            with F.Context(_params(shape), storage=storage, lane_cells=lc) as ctx:
                assert ctx.launch_shape()["tile_x"] == 64 * lc
                ctx.set_materials(ids, table)
                ctx.set_sources(src, jz)
                ctx.run(9)
                out.append([ctx.get(f, lv) for f in range(9) for lv in ([0] if 3 <= f < 6 else [0, 1, 2])])
        for a, b in zip(*out):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), shape
        assert np.abs(out[0][2]).max() > 0


def test_face_sources_batch_equals_single():
    from tests.test_gpu_batch import make_case
    insts = []
    for q, shape in enumerate([(37, 29, 23), (65, 20, 30)]):
        x = make_case(800 + q, shape, 5, 12, storage=F.STORE_F32 if q == 0 else F.STORE_F16)
        x.src = face_sources(shape)
        x.c.nsrc = x.src.shape[0]
        x.c.src_xyz = F._ptr(x.src)
        insts.append(x)
    refs = [single(x) for x in insts]
    with F.Batch(1 << 28) as b:
        ids = [b.add(x) for x in insts]
        b.run(12)
        for i, x, (fields, series) in zip(ids, insts, refs):
            for (f, lv), want in fields.items():
                assert np.array_equal(b.get(i, f, lv).view(np.uint32), want.view(np.uint32)), (x.shape, f, lv)
            assert np.array_equal(b.probes(i).view(np.uint64), series.view(np.uint64))


def test_face_source_from_the_cli(tmp_path):
    """A parameter file whose source sits on the x-high face (1-based x = nx)
    runs on both CLI paths with the same output."""
    from tests.test_gpu_cli import SHAPE, _case, _run
    params, text = _case(tmp_path, t_max=3)
    nx, ny, nz = SHAPE
    text = text.replace("source_location: 12 10 8", f"source_location: {nx} {ny // 2} {nz}")
    fused = _run(params, text, tmp_path / "out_fused", "--lines")
    legacy = _run(params, text, tmp_path / "out_legacy", "--lines", "--legacy")
    assert fused == legacy and len(fused) == 9
    vals = [float(v) for text_ in fused.values() for line in text_.split("\n") if line
            for v in line.split()[3:]]
    assert any(v != 0.0 for v in vals)


def test_sources_outside_the_addressable_cells_are_refused():
    shape = (20, 18, 16)
    with F.Context(_params(shape)) as ctx:
        for bad in ([0, 5, 5], [5, 0, 5], [20, 5, 5], [5, 18, 5], [5, 5, 16], [5, 5, -1]):
            with pytest.raises(F.FdtdError, match="-2"):
                ctx.set_sources(np.asarray([bad], np.int32), np.ones(4, np.float32))
        ctx.set_sources(np.asarray([[19, 17, 15]], np.int32), np.ones(4, np.float32))
