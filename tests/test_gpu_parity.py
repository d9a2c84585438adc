"""GPU parity of the HIP path against the CPU oracle's FP32 transcription.

Every HIP update reproduces the reference's operand order and precision
(float products/sums, the double Debye sums, correctly rounded sqrt and
division, no contraction), so the bar here is bit equality of every field,
slot and probe sample with the oracle on the same seeded inputs."""
import ctypes as C

import numpy as np
import pytest

import fdtd_amd as F
from fdtd_amd import hip
from fdtd_amd import workloads as W
from oracle import oracle as O
from tests.helpers import (EPS0, MU0, bits_equal, first_diff, oracle_sim, random_materials, random_rp,
                           random_sources, random_state, rel_l2)

pytestmark = pytest.mark.gpu

DT = 3.8105118e-12
DX = 2.0e-3


def _params(shape):
    return F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0)


class Legacy:
    """The reference's FdtdField layout driven through the five drop-in
    entry points, step by step as fdtd.cu:115-530 does."""

    def __init__(self, shape, table, ids, rps, sources, jz, debye):
        self.shape = shape
        nx, ny, nz = shape
        n = nx * ny * nz
        self.E = [[hip.DeviceArray(n) for _ in range(3)] for _ in range(3)]
        self.D = [[hip.DeviceArray(n) for _ in range(3)] for _ in range(3)]
        self.H = [hip.DeviceArray(n) for _ in range(3)]
        t = table[ids.astype(np.int64)]
        self.mat = [hip.DeviceArray.from_host(t[..., q].ravel()) for q in (0, 2, 1, 3)]  # sigma, epsI, epsS, tauD
        self.rp = [hip.DeviceArray.from_host(r) for r in rps]
        p = _params(shape)
        self._src = (C.c_int * max(1, 3 * len(sources)))(*np.asarray(sources, np.int32).ravel().tolist())
        self._jz = (C.c_float * max(1, len(jz)))(*np.asarray(jz, np.float32).tolist())
        p.sourcesCount = len(sources)
        p.sources = C.cast(self._src, C.POINTER(C.c_int))
        p.jzCount = len(jz)
        p.jz = C.cast(self._jz, C.POINTER(C.c_float))
        L = F.lib()
        F.check(L.copySymbolsToDevice(C.byref(p)), "copySymbolsToDevice")
        F.check(L.fdtdSetDebyeMode(debye), "fdtdSetDebyeMode")
        self.n = 0

    def set_state(self, st):
        for (f, lv), a in st.items():
            self.dev(f, lv).upload(a.ravel())

    def dev(self, f, lv=0):
        q = f % 3
        if 3 <= f < 6:
            return self.H[q]
        s = (self.n + 2 - lv) % 3
        return self.D[s][q] if f < 3 else self.E[s][q]

    def step(self, jz_index):
        L = F.lib()
        n = self.n
        t, s1, s2 = n % 3, (n + 2) % 3, (n + 1) % 3
        E, D, H = self.E, self.D, self.H
        p = lambda arrs: [a.ptr for a in arrs]
        F.check(L.updateHField(*p(H), *p(E[s1]), None), "updateHField")
        F.check(L.updateDField(*p(D[t]), *p(D[s1]), *p(H), None), "updateDField")
        F.check(L.updateSources(D[t][2].ptr, D[s1][2].ptr, H[0].ptr, H[1].ptr, jz_index, None), "updateSources")
        F.check(L.updateEField(*p(E[t]), *p(E[s1]), *p(E[s2]), *p(D[t]), *p(D[s1]), *p(D[s2]), *p(self.mat), None),
                "updateEField")
        F.check(L.updateMurBoundary(*p(E[t]), *p(E[s1]), *p(self.rp), None), "updateMurBoundary")
        self.n = n + 1

    def get(self, f, lv=0):
        hip.sync()
        nx, ny, nz = self.shape
        return self.dev(f, lv).download().reshape(nz, ny, nx)


def _compare_all(get_a, get_b, tag):
    for f in range(9):
        for lv in ([0] if 3 <= f < 6 else [0, 1, 2]):
            a, b = get_a(f, lv), get_b(f, lv)
            assert bits_equal(a, b), f"{tag}: {O.FIELDS[f]} level {lv}: {first_diff(a, b)}"


@pytest.mark.parametrize("debye", [O.DEBYE_FIXED, O.DEBYE_REF])
def test_legacy_entry_points_bit_exact(debye):
    """Each drop-in entry point, random state, odd grid: bit equality with
    the oracle kernels after 6 full steps (sources held for 3 steps)."""
    print(f"legacy debye={debye}", flush=True)
    rng = np.random.default_rng(7 + debye)
    shape = (37, 29, 23)
    ids, table = random_materials(rng, shape)
    rps = random_rp(rng, shape)
    src = random_sources(rng, shape, 5)
    jz = rng.uniform(-2, 2, 40).astype(np.float32)
    st = random_state(rng, shape)
    leg = Legacy(shape, table, ids, rps, src, jz, debye)
    leg.set_state(st)
    sim = oracle_sim(shape, DT, DX, debye=debye)
    t = table[ids.astype(np.int64)]
    sim.set_materials(t[..., 0], t[..., 1], t[..., 2], t[..., 3])
    sim.set_rp(rps)
    sim.set_sources(src, jz, jz_mode=0)
    for (f, lv), a in st.items():
        sim.set(f, lv, a)
    for n in range(6):
        leg.step((n // 3) * 3)
        sim.step()
    _compare_all(leg.get, sim.get, "legacy")


FUSED_CASES = [
    # shape, tile_y, kchunk, debye, jz mode, threads per tile row
    ((37, 29, 23), 4, 0, O.DEBYE_FIXED, F.JZ_STEP, 64),
    ((37, 29, 23), 2, 5, O.DEBYE_REF, F.JZ_HOLD3, 64),
    ((64, 64, 64), 8, 7, O.DEBYE_FIXED, F.JZ_STEP, 64),
    ((130, 17, 9), 2, 3, O.DEBYE_FIXED, F.JZ_HOLD3, 64),
    ((5, 6, 7), 4, 0, O.DEBYE_FIXED, F.JZ_STEP, 64),
    ((65, 5, 40), 4, 40, O.DEBYE_REF, F.JZ_STEP, 64),
    ((130, 17, 9), 4, 3, O.DEBYE_FIXED, F.JZ_STEP, 128),
    ((200, 11, 13), 8, 4, O.DEBYE_REF, F.JZ_HOLD3, 128),
    ((65, 9, 12), 2, 0, O.DEBYE_FIXED, F.JZ_STEP, 128),
    ((129, 7, 10), 4, 16, O.DEBYE_FIXED, F.JZ_STEP, 128),
    ((37, 29, 23), 4, 0, O.DEBYE_FIXED, F.JZ_STEP, 128),
    # x-high face on the first cell of a wave (nx-1 a multiple of 64): the
    # x-high Mur faces stay in the Mur kernel; tall tiles put most rows' face
    # neighbour in another wave's staged column
    ((65, 20, 30), 8, 0, O.DEBYE_FIXED, F.JZ_STEP, 64),
    ((129, 9, 12), 4, 5, O.DEBYE_REF, F.JZ_STEP, 64),
    ((193, 9, 12), 8, 4, O.DEBYE_FIXED, F.JZ_STEP, 128),
    # face inside the second wave of a 128-wide tile (folded into the tile)
    ((100, 13, 11), 8, 0, O.DEBYE_FIXED, F.JZ_HOLD3, 128),
    # the smallest grids: every edge 3 (no interior in that direction beyond one cell)
    ((3, 3, 3), 2, 0, O.DEBYE_FIXED, F.JZ_STEP, 64),
    ((3, 9, 5), 4, 2, O.DEBYE_REF, F.JZ_STEP, 64),
    ((70, 3, 4), 8, 0, O.DEBYE_FIXED, F.JZ_HOLD3, 64),
    ((66, 7, 3), 2, 1, O.DEBYE_FIXED, F.JZ_STEP, 128),
    # rows padded by one line past a multiple of 2 KiB (pitch 1056 for both)
    ((1024, 9, 10), 4, 0, O.DEBYE_FIXED, F.JZ_STEP, 64),
    ((1000, 7, 9), 2, 5, O.DEBYE_REF, F.JZ_HOLD3, 128),
]


@pytest.mark.parametrize("shape,tile_y,kchunk,debye,jz_mode,tile_x", FUSED_CASES)
def test_fused_step_bit_exact(shape, tile_y, kchunk, debye, jz_mode, tile_x):
    """Fused H/D/E kernel + Mur faces over random state, partial tiles,
    several launch shapes: every field of every slot bit-equal."""
    print(f"fused {shape} ty={tile_y} kc={kchunk} debye={debye} jz={jz_mode} tx={tile_x}", flush=True)
    rng = np.random.default_rng(sum(shape) + tile_y)
    ids, table = random_materials(rng, shape)
    rps = random_rp(rng, shape)
    src = random_sources(rng, shape, 4)
    jz = rng.uniform(-2, 2, 20).astype(np.float32)
    st = random_state(rng, shape)
    ctx = F.Context(_params(shape), debye=debye, jz_mode=jz_mode, tile_y=tile_y, kchunk=kchunk, tile_x=tile_x)
    assert ctx.launch_shape()["tile_x"] == tile_x
    ctx.set_materials(ids, table)
    ctx.set_rp(rps)
    ctx.set_sources(src, jz)
    for (f, lv), a in st.items():
        ctx.set(f, lv, a)
    sim = oracle_sim(shape, DT, DX, debye=debye)
    t = table[ids.astype(np.int64)]
    sim.set_materials(t[..., 0], t[..., 1], t[..., 2], t[..., 3])
    sim.set_rp(rps)
    sim.set_sources(src, jz, jz_mode=0 if jz_mode == F.JZ_HOLD3 else 1)
    for (f, lv), a in st.items():
        sim.set(f, lv, a)
    for _ in range(7):
        ctx.step()
        sim.step()
    _compare_all(ctx.get, sim.get, f"fused {shape}")
    ctx.close()


def test_plain_small_probe_series():
    """64^3 Debye sphere, 500 steps: the Ez probe series and the final
    fields are bit-equal to the FP32 transcription; the error against the
    FP64 oracle is reported."""
    print("plain_small 500 steps", flush=True)
    w = W.plain_small()
    ctx = F.Context(w.params)
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, w.steps)
    ctx.run(w.steps)
    got = ctx.probes()
    sims = {}
    for prec in (32, 64):
        s = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0, prec=prec)
        s.set_materials(*w.material_arrays())
        s.set_sources(w.sources, w.jz, jz_mode=1)
        sims[prec] = (s, s.run(w.steps, w.probes))
    ref32 = sims[32][1]
    assert bits_equal(got, ref32.astype(np.float32)), first_diff(got, ref32)
    for f in range(9):
        assert bits_equal(ctx.get(f), sims[32][0].get(f)), O.FIELDS[f]
    err = rel_l2(got, sims[64][1])
    print(f"plain_small probe rel-L2 vs FP64 oracle: {err:.3e}")
    assert err < 1e-4
    assert np.abs(got).max() > 0


def test_plain_256_prefix_bit_exact():
    """256^3 eight-ellipsoid phantom with the 64k-cell Jz sheet: 30 steps
    bit-equal to the oracle (all nine fields), probes and z-slices included."""
    print("plain_256 30 steps", flush=True)
    w = W.plain_256(steps=30)
    ctx = F.Context(w.params)
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, w.steps)
    ctx.set_slices(w.slice_k, 10, 8)
    ctx.run(w.steps)
    sim = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0, prec=32)
    sim.set_materials(*w.material_arrays())
    sim.set_sources(w.sources, w.jz, jz_mode=1)
    ref = sim.run(w.steps, w.probes)
    got = ctx.probes()
    assert bits_equal(got, ref.astype(np.float32)), first_diff(got, ref)
    for f in range(9):
        a, b = ctx.get(f), sim.get(f)
        assert bits_equal(a, b), f"{O.FIELDS[f]}: {first_diff(a, b)}"
    steps, data = ctx.slices()
    assert list(steps) == [10, 20, 30]
    last = data[-1]
    for f in range(9):
        assert bits_equal(last[f], ctx.get(f)[w.slice_k]), O.FIELDS[f]


def test_probes_match_fields_each_step():
    """Probe rows are the newest fields at the probe cells after each step."""
    rng = np.random.default_rng(3)
    shape = (40, 33, 21)
    ids, table = random_materials(rng, shape)
    probes = np.asarray([[f, 3 + f, 5 + f, 2 + f] for f in range(9)], np.int32)
    ctx = F.Context(_params(shape))
    ctx.set_materials(ids, table)
    ctx.set_sources(random_sources(rng, shape, 3), rng.uniform(-1, 1, 10).astype(np.float32))
    ctx.set_probes(probes, 5)
    for n in range(5):
        ctx.step()
        row = ctx.probes()[n]
        for p, (f, i, j, k) in enumerate(probes):
            assert row[p] == np.float64(ctx.get(f)[k, j, i])


def _face_probes(shape):
    """One cell of each of the 12 Mur segments, E cells on faces, edges and
    corners that no segment updates, and D/H cells on faces."""
    nx, ny, nz = shape
    i, j, k = nx // 2, ny // 3, nz // 2
    cells = [
        (6, i, 0, k), (6, i, ny - 1, k), (6, i, j, 0), (6, i, j, nz - 1),          # Ex segments 0-3
        (7, 0, j, k), (7, nx - 1, j, k), (7, i, j, 0), (7, i, j, nz - 1),          # Ey segments 4-7
        (8, 0, j, k), (8, nx - 1, j, k), (8, i, 0, k), (8, i, ny - 1, k),          # Ez segments 8-11
        (6, 0, 0, 0), (6, nx - 1, 0, k), (6, i, 0, 0), (7, 0, ny - 1, k), (7, 0, j, 0),
        (8, 0, 0, k), (8, i, j, nz - 1), (8, nx - 1, ny - 1, nz - 1), (7, nx - 1, ny - 2, nz - 2),
        (2, 0, j, k), (3, i, 0, k), (5, i, j, nz - 1), (1, nx - 1, j, 0),
    ]
    return np.asarray(cells, np.int32)


@pytest.mark.parametrize("shape,kchunk,tile_x", [((40, 21, 19), 4, 64),    # x faces in the fused kernel
                                                 ((65, 12, 11), 4, 64),    # x-high faces in the Mur kernel
                                                 ((65, 12, 11), 3, 128),
                                                 ((40, 21, 19), 20, 64)])  # all faces in the Mur kernel
def test_face_probes_match_fields(shape, kchunk, tile_x):
    """Probe rows are recorded in the Mur launch; on an E face cell that
    launch updates, the probe evaluates the face itself.  Every row equals
    the fields read back after the step."""
    print(f"face probes {shape} kc={kchunk} tx={tile_x}", flush=True)
    rng = np.random.default_rng(11)
    ids, table = random_materials(rng, shape)
    probes = _face_probes(shape)
    ctx = F.Context(_params(shape), tile_y=4, kchunk=kchunk, tile_x=tile_x)
    ctx.set_materials(ids, table)
    ctx.set_rp(random_rp(rng, shape))
    ctx.set_sources(random_sources(rng, shape, 2), rng.uniform(-1, 1, 10).astype(np.float32))
    for (f, lv), a in random_state(rng, shape).items():
        ctx.set(f, lv, a)
    steps = 6
    ctx.set_probes(probes, steps)
    for n in range(steps):
        ctx.step()
        row = ctx.probes()[n]
        for p, (f, i, j, k) in enumerate(probes):
            want = ctx.get(f)[k, j, i]
            assert want != 0.0 or p >= 12, f"probe {p}: segment face value is zero"
            assert row[p] == np.float64(want), f"step {n + 1} probe {p} {O.FIELDS[f]} ({i},{j},{k})"


@pytest.mark.parametrize("storage,nx,pitch", [(F.STORE_F32, 1024, 1056), (F.STORE_F32, 1000, 1056),
                                              (F.STORE_F32, 1536, 1568), (F.STORE_F32, 512, 512),
                                              (F.STORE_F32, 256, 256), (F.STORE_F32, 37, 64),
                                              (F.STORE_F16, 2048, 2112), (F.STORE_F16, 1024, 1024)])
def test_row_pitch(storage, nx, pitch):
    """Rows start on 128-byte lines; rows of 4 KiB or more whose length is a
    multiple of 2 KiB take one more line."""
    with F.Context(_params((nx, 4, 4)), storage=storage) as ctx:
        assert ctx.layout()[0] == pitch
