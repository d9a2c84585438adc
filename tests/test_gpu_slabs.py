"""z-slab decomposition on one GPU: a grid cut into slabs along z (each with
ghost planes toward its neighbours) and stepped with one halo exchange per
interface per step (fdtd_ctx_halo_pack / fdtd_ctx_halo_unpack through device
message buffers) equals the whole grid bit for bit on every real plane, every
field and every ring level."""
import numpy as np
import pytest

import fdtd_amd as F
from fdtd_amd import hip
from tests.helpers import EPS0, MU0, random_materials, random_rp, random_sources, random_state

pytestmark = pytest.mark.gpu

DT = 3.8105118e-12
DX = 2.0e-3


def _slab_planes(nz, cuts):
    """(first local plane, end, first real, end real) in global planes, per slab."""
    b = [0] + list(cuts) + [nz]
    out = []
    for s in range(len(b) - 1):
        lo, hi = b[s], b[s + 1]
        out.append((lo - (s > 0), hi + (s < len(b) - 2), lo, hi))
    return out


def _slice_rp(rps, shape, g0, g1):
    nx, ny, nz = shape
    rx = [r.reshape(nz, ny, 2)[g0:g1].ravel() for r in (rps[0], rps[3])]
    ry = [r.reshape(nz, 2, nx)[g0:g1].ravel() for r in (rps[1], rps[4])]
    return [rx[0], ry[0], rps[2], rx[1], ry[1], rps[5]]


@pytest.mark.parametrize("storage,cuts", [(F.STORE_F32, (19,)), (F.STORE_F32, (13, 27)), (F.STORE_F16, (13, 27))])
def test_slabs_equal_whole_grid(storage, cuts):
    shape = (37, 29, 40)
    nx, ny, nz = shape
    print(f"slabs {shape} cuts={cuts} storage={storage}", flush=True)
    rng = np.random.default_rng(23)
    ids, table = random_materials(rng, shape)
    rps = random_rp(rng, shape)
    src = random_sources(rng, shape, 6)
    src[0, 2], src[1, 2] = cuts[0] - 1, cuts[0]  # sources on both sides of the first cut
    jz = rng.uniform(-2, 2, 30).astype(np.float32)
    # random fields for fp32; fp16 starts from zero (sources on both sides of
    # the first cut carry the fields across it)
    st = random_state(rng, shape) if storage == F.STORE_F32 else {}

    def ctx_for(sh, rp, sources, planes):
        c = F.Context(F.make_params(sh, DT, DX, mu0=MU0, eps0=EPS0), storage=storage)
        c.set_materials(ids[planes], table)
        c.set_rp(rp)
        c.set_sources(sources, jz)
        for (f, lv), a in st.items():
            c.set(f, lv, a[planes])
        return c

    whole = ctx_for(shape, rps, src, slice(0, nz))
    slabs = []
    for g0, g1, r0, r1 in _slab_planes(nz, cuts):
        own = src[(src[:, 2] >= r0) & (src[:, 2] < r1)].copy()
        own[:, 2] -= g0
        slabs.append((ctx_for((nx, ny, g1 - g0), _slice_rp(rps, shape, g0, g1), own, slice(g0, g1)), g0, r0, r1))
    up = hip.DeviceArray(slabs[0][0].halo_bytes(F.HALO_UP) // 4)
    down = hip.DeviceArray(slabs[0][0].halo_bytes(F.HALO_DOWN) // 4)
    for _ in range(9):
        for (lo, _, _, _), (hi, _, _, _) in zip(slabs, slabs[1:]):
            lo.halo_pack(F.HALO_UP, up.ptr)
            hi.halo_unpack(F.HALO_UP, up.ptr)
            hi.halo_pack(F.HALO_DOWN, down.ptr)
            lo.halo_unpack(F.HALO_DOWN, down.ptr)
        for c, _, _, _ in slabs:
            c.step()
        whole.step()
    for f in range(9):
        for lv in ([0] if 3 <= f < 6 else [0, 1, 2]):
            ref = whole.get(f, lv)
            for c, g0, r0, r1 in slabs:
                got = c.get(f, lv)[r0 - g0:r1 - g0]
                want = ref[r0:r1]
                bad = np.argwhere(got.view(np.uint32) != want.view(np.uint32))
                assert bad.size == 0, f"field {f} level {lv} slab planes [{r0},{r1}): {len(bad)} cells, first {bad[0]}"
    up.free()
    down.free()


@pytest.mark.parametrize("cuts", [(30,), (20, 33, 45)])
def test_slabs_with_zero_term_rows(cuts):
    """From zero fields (no set_field), the slabs keep the zero-term rows across
    halo exchanges (a vacuum grid with a Debye sphere: most rows skip their
    E^{n-1}/D^{n-1} loads) and still equal the whole grid bit for bit, with
    the source's wave crossing every cut."""
    from fdtd_amd import workloads as W
    w = W.plain_small(steps=80)
    nx, ny, nz = w.shape
    print(f"zero-term slabs {w.shape} cuts={cuts}", flush=True)

    def ctx_for(sh, sources, planes):
        c = F.Context(F.make_params(sh, w.params.dt, w.params.dx, mu0=w.params.mu0, eps0=w.params.eps0))
        c.set_materials(w.ids[planes], w.table)
        c.set_sources(sources, w.jz)
        return c

    whole = ctx_for(w.shape, w.sources, slice(0, nz))
    slabs = []
    for g0, g1, r0, r1 in _slab_planes(nz, cuts):
        own = w.sources[(w.sources[:, 2] >= r0) & (w.sources[:, 2] < r1)].copy()
        own[:, 2] -= g0
        slabs.append((ctx_for((nx, ny, g1 - g0), own, slice(g0, g1)), g0, r0, r1))
    up = hip.DeviceArray(slabs[0][0].halo_bytes(F.HALO_UP) // 4)
    down = hip.DeviceArray(slabs[0][0].halo_bytes(F.HALO_DOWN) // 4)
    for _ in range(80):
        for (lo, _, _, _), (hi, _, _, _) in zip(slabs, slabs[1:]):
            lo.halo_pack(F.HALO_UP, up.ptr)
            hi.halo_unpack(F.HALO_UP, up.ptr)
            hi.halo_pack(F.HALO_DOWN, down.ptr)
            lo.halo_unpack(F.HALO_DOWN, down.ptr)
        for c, _, _, _ in slabs:
            c.step()
        whole.step()
    for f in range(9):
        for lv in ([0] if 3 <= f < 6 else [0, 1, 2]):
            ref = whole.get(f, lv)
            for c, g0, r0, r1 in slabs:
                got = c.get(f, lv)[r0 - g0:r1 - g0]
                want = ref[r0:r1]
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"field {f} level {lv} [{r0},{r1})"
    # the wave reached every slab
    ez = whole.get(8)
    for _, _, r0, r1 in slabs:
        assert np.abs(ez[r0:r1]).max() > 0
    up.free()
    down.free()
