"""16-bit storage with two cells per lane (one 32-bit access per stream)
against one cell per lane: the per-cell arithmetic is the same operation for
operation, so every field and probe must agree bit for bit.  Shapes cover
odd widths, widths below one 128-cell tile, x-high faces on either cell of
a pair, the Mur fallback when nx-1 is a multiple of 128, sources, probes and
several z-chunk lengths."""
import numpy as np
import pytest

import fdtd_amd as F
from tests.helpers import EPS0, MU0, first_diff, random_materials, random_rp, random_sources

pytestmark = pytest.mark.gpu

DT = 3.8105118e-12
DX = 2.0e-3


def _jz(n=60):
    t = np.arange(n, dtype=np.float64)
    return (-(t - 12) / 4 * np.exp(-((t - 12) / 4) ** 2) * 2.0).astype(np.float32)


def run(shape, storage, lane_cells, kchunk, steps, seed):
    rng = np.random.default_rng(seed)
    ids, table = random_materials(rng, shape, 6)
    ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0), storage=storage, lane_cells=lane_cells,
                    kchunk=kchunk)
    ctx.set_materials(ids, table)
    ctx.set_rp(random_rp(rng, shape))
    ctx.set_sources(random_sources(rng, shape, 3), _jz())
    nx, ny, nz = shape
    probes = np.asarray([[8, nx // 2, ny // 2, nz // 2], [7, 0, ny // 2, nz // 2], [8, nx - 1, ny // 3, 1],
                         [2, 1, 1, 0], [4, nx - 2, ny - 2, nz - 2]], np.int32)
    ctx.set_probes(probes, steps)
    ctx.run(steps)
    fields = {(f, lv): ctx.get(f, lv) for f in range(9) for lv in ([0] if 3 <= f < 6 else [0, 1, 2])}
    return fields, ctx.probes(), ctx.launch_shape()


@pytest.mark.parametrize("storage", [F.STORE_F16, F.STORE_BF16])
@pytest.mark.parametrize("shape,kchunk", [((37, 29, 23), 0), ((130, 17, 9), 4), ((129, 12, 14), 8), ((128, 9, 10), 3),
                                          ((257, 6, 7), 0), ((64, 33, 40), 16), ((5, 6, 7), 0), ((200, 20, 30), 5),
                                          ((2048, 6, 7), 0)])   # 4 KiB rows, padded by a line
def test_pairs_equal_single_cells(storage, shape, kchunk):
    print(f"\n[pairs] {shape} kc={kchunk} storage={storage}", flush=True)
    a, pa, sa = run(shape, storage, 2, kchunk, 25, 71)
    b, pb, sb = run(shape, storage, 1, kchunk, 25, 71)
    assert sa["tile_x"] == 128 and sb["tile_x"] == 64
    for key in a:
        assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), f"{F.FIELDS[key[0]]} level {key[1]}: {first_diff(a[key], b[key])}"
    assert np.array_equal(pa.view(np.uint64), pb.view(np.uint64))
    assert np.isfinite(pa).all()


@pytest.mark.parametrize("storage", [F.STORE_F32, F.STORE_F16])
def test_ctx_autotune_restores_state(storage):
    shape = (70, 40, 33)
    rng = np.random.default_rng(9)
    ids, table = random_materials(rng, shape, 5)
    src = random_sources(rng, shape, 2)

    def make():
        ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0), storage=storage)
        ctx.set_materials(ids, table)
        ctx.set_sources(src, _jz())
        ctx.set_probes(np.asarray([[8, 35, 20, 16]], np.int32), 30)
        return ctx

    a = make()
    ty, kc, lc = a.autotune(trial_steps=2)
    assert ty in (2, 4, 8) and kc in (3, 4, 6, 8, 16) and lc in ((1, 2) if storage == F.STORE_F16 else (1,))
    assert a.n == 0
    a.run(30)
    b = make()
    b.run(30)
    for f in range(9):
        assert np.array_equal(a.get(f).view(np.uint32), b.get(f).view(np.uint32)), F.FIELDS[f]
    assert np.array_equal(a.probes().view(np.uint64), b.probes().view(np.uint64))
    with pytest.raises(F.FdtdError, match="-3"):
        a.autotune()


@pytest.mark.parametrize("storage,shape,tile_x,kchunk", [
    (F.STORE_F16, (256, 256, 256), 128, 16), (F.STORE_BF16, (320, 320, 880), 64, 16), (F.STORE_F16, (96, 96, 96), 64, 4),
    (F.STORE_F16, (200, 100, 100), 64, 4), (F.STORE_BF16, (128, 128, 128), 128, 4), (F.STORE_F16, (64, 64, 64), 64, 4),
    (F.STORE_F32, (320, 320, 880), 64, 16), (F.STORE_F32, (256, 256, 256), 64, 8), (F.STORE_F32, (192, 192, 192), 64, 8),
    (F.STORE_F32, (96, 96, 96), 64, 4)])
def test_default_launch_shapes(storage, shape, tile_x, kchunk):
    """Default shapes of a context (4 tile rows): 16-bit grids take pairs where
    128-cell rows are at least 90% full and a pair launch fills the GPU's
    resident blocks; the z-chunk is the one of least cost in rounds of
    resident blocks times (planes + 2)."""
    with F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0), storage=storage) as ctx:
        s = ctx.launch_shape()
        assert (s["tile_x"], s["tile_y"], s["kchunk"]) == (tile_x, 4, kchunk), s
