"""Seeded synthetic workloads of the FDTD path.

Every random draw comes from an in-house splitmix64 stream with an explicit
float mapping (53 high-quality bits -> [0,1)), so the same seed builds the
same grids on any host.  Grid parameters are derived by the library's own
initParamsWithPath from a parameter file in the reference's format
(data/input_params) plus extension keys, and the source waveform by its
setupSources (fdtd.cu:1027-1082), so the workloads exercise the same setup
code as the CLI.

Workloads (sizes as stated):
  plain_small   64^3, 2 mm cells, Courant 0.99, vacuum plus a centred Debye
                sphere of radius 12 (eps_inf 40, eps_s 55, tau 8 ps,
                sigma 0.8 S/m), Ez point source at (16,32,32), 2.45 GHz,
                500 steps, Ez probe at the centre.
  plain_256     256^3, 8 nested seeded ellipsoids (semi-axes 20-110 cells,
                centres 128 +- 10), random Debye materials, Jz sheet on the
                plane x = 10, 900 MHz, 2000 steps, 3 probes, z-slice k = 128
                every 100 steps.
  large_phantom 320x320x880, 50 tissue ids from seeded ellipsoids and
                z-cylinders, 10-cell Jz dipole at the phantom edge, 900 MHz,
                5000 steps.
  batched_set   64 independent grids, edges nx, ny in [48,160] and nz in
                [48,224] (48^3 to 160x160x224), each with its own phantom of
                4-12 Debye materials (nested ellipsoids), a point source at a
                seeded position, a carrier in [300 MHz, 6 GHz] and 500-3000
                steps; fp32.
  riken         the RIKEN voxel phantom (265x490x1687, 53 tissues, tissue
                table data/mat_specs_riken) padded by 16 vacuum cells per
                side, with the reference's own source settings
                (data/input_params: 3 GHz, pulse width 0.5, no carrier
                modulation, 100 cells per wavelength, so 1 mm cells) at
                Courant 0.99, a soft Jz source at the grid centre.
  box_phantom   any grid: vacuum with four seeded axis-aligned Debye boxes
                kept PHANTOM_MARGIN cells inside the faces, a centred Ez point
                source at 900 MHz and a centre probe; cheap to build at
                billions of cells.
  served_stream Poisson arrivals (mean 40 per second) of solves with edges
                in [32,192] per axis, 200-4000 steps, fp32 or fp16 storage
                (even odds), 4-12 material ellipsoid phantoms and a point or
                a z-dipole source (even odds).
"""
import ctypes as C
import dataclasses
import os
import tempfile

import numpy as np

from . import FdtdParams, lib

C0 = 3.0e8


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with a 53-bit float mapping."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = seed & self.MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """Integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


@dataclasses.dataclass
class Workload:
    name: str
    shape: tuple            # (nx, ny, nz)
    params: FdtdParams
    ids: np.ndarray         # uint8 (nz, ny, nx)
    table: np.ndarray       # float32 (nmat, 4): sigma, eps_s, eps_inf, tau
    sources: np.ndarray     # int32 (nsrc, 3), 0-based
    jz: np.ndarray          # float32 waveform, one entry per step
    probes: np.ndarray      # int32 (nprobe, 4): field id, i, j, k
    steps: int
    slice_k: int = -1
    slice_every: int = 0
    seed: int = 0
    freq: float = 0.0

    @property
    def cells(self):
        nx, ny, nz = self.shape
        return nx * ny * nz

    def material_arrays(self):
        """The reference's four per-voxel arrays (fdtd.h:50-51)."""
        t = self.table[self.ids.astype(np.int64)]
        return t[..., 0].copy(), t[..., 1].copy(), t[..., 2].copy(), t[..., 3].copy()

    def rp_planes(self):
        """Mur permittivity planes of the vacuum default material: 1.0."""
        nx, ny, nz = self.shape
        return [np.ones(n, np.float32) for n in (2 * ny * nz, nx * 2 * nz, nx * ny * 2) * 2]


def params_text(shape, steps, freq, cell, courant, pulse_width=0.5, pmf=1.0, sources=((1, 1, 1),)):
    """A parameter file in the reference's positional format
    (data/input_params) followed by extension keys."""
    nx, ny, nz = shape
    lines = [
        f"nx_ny_nz: {nx} {ny} {nz}", f"t_max: {steps}", "nf: 1", "env_set_dir: data/pgmdata",
        'env_file_prefix: "ver"', "output_dir: output", 'output_format: "txt"', "impulse_resp_flag: 0",
        "pec_flag: 0", "read_env_flag: 0", "output_flag: 1", "bzip2_flag: 0", "output_start: 0 0 0",
        "output_finish: 0 0 0", "source_type:  soft", "elements_per_wavelength: 100",
        f"wave_frequency: {freq:.9e}", f"pulse_width: {pulse_width}", f"pulse_modulation_frequency: {pmf}",
        f"number_of_excitation_sources: {len(sources)}",
    ]
    lines += [f"source_location: {s[0] + 1} {s[1] + 1} {s[2] + 1}" for s in sources]
    lines += ["pulse_type: 2", "fsigma: 0", "feps_s: 1", "feps_inf: 1", "ftau_d: 0",
              f"courant_factor: {courant}", f"cell_size: {cell:.9e}"]
    return "\n".join(lines) + "\n"


def derive(shape, steps, freq, cell, courant, pulse_width=0.5, pmf=1.0):
    """(FdtdParams copy, jz waveform) from the library's own
    initParamsWithPath + setupSources."""
    L = lib()
    with tempfile.NamedTemporaryFile("w", suffix=".params", delete=False) as f:
        f.write(params_text(shape, steps, freq, cell, courant, pulse_width, pmf))
        path = f.name
    try:
        pp = L.initParamsWithPath(path.encode())
        L.setupSources(pp)
        p = pp.contents
        jz = np.ctypeslib.as_array(p.jz, shape=(p.jzCount,)).astype(np.float32).copy()
        out = FdtdParams()
        C.memmove(C.byref(out), C.byref(p), C.sizeof(FdtdParams))
        out.inputPath = None
        out.outputPath = None
        out.sources = None
        out.jz = None
        out.sourcesCount = 0
        out.jzCount = 0
        L.deallocParams(pp)
    finally:
        os.unlink(path)
    return out, jz


def debye_table(rng, n):
    """n random Debye materials: eps_inf U[4,45], eps_s U[5,80] (the pair
    sorted so eps_s >= eps_inf), tau U[5,50] ps, sigma U[0.02,2.5] S/m."""
    rows = []
    for _ in range(n):
        ei = rng.uniform(4.0, 45.0)
        es = rng.uniform(5.0, 80.0)
        ei, es = min(ei, es), max(ei, es)
        tau = rng.uniform(5e-12, 50e-12)
        sg = rng.uniform(0.02, 2.5)
        rows.append((sg, es, ei, tau))
    return np.asarray(rows, np.float32)


def _paint_ellipsoid(ids, c, a, value):
    nz, ny, nx = ids.shape
    lo = [max(0, int(np.floor(c[d] - a[d]))) for d in range(3)]
    hi = [min(n, int(np.ceil(c[d] + a[d])) + 1) for d, n in zip(range(3), (nx, ny, nz))]
    if any(h <= l for l, h in zip(lo, hi)):
        return
    x = (np.arange(lo[0], hi[0]) - c[0]) / a[0]
    y = (np.arange(lo[1], hi[1]) - c[1]) / a[1]
    z = (np.arange(lo[2], hi[2]) - c[2]) / a[2]
    m = (z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2) <= 1.0
    sub = ids[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    sub[m] = value


def _paint_zcylinder(ids, c, r, z0, z1, value):
    nz, ny, nx = ids.shape
    lo = [max(0, int(np.floor(c[0] - r))), max(0, int(np.floor(c[1] - r))), max(0, z0)]
    hi = [min(nx, int(np.ceil(c[0] + r)) + 1), min(ny, int(np.ceil(c[1] + r)) + 1), min(nz, z1)]
    if any(h <= l for l, h in zip(lo, hi)):
        return
    x = np.arange(lo[0], hi[0]) - c[0]
    y = np.arange(lo[1], hi[1]) - c[1]
    m2 = (y[:, None] ** 2 + x[None, :] ** 2) <= r * r
    sub = ids[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    sub[:, m2] = value


def plain_small(steps=500):
    """64^3 vacuum with a Debye sphere; CPU-runnable reference case."""
    shape = (64, 64, 64)
    params, jz = derive(shape, steps, 2.45e9, 2.0e-3, 0.99)
    ids = np.zeros(shape[::-1], np.uint8)
    k, j, i = np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij")
    ids[(i - 32) ** 2 + (j - 32) ** 2 + (k - 32) ** 2 <= 144] = 1
    table = np.asarray([(0.0, 1.0, 1.0, 0.0), (0.8, 55.0, 40.0, 8e-12)], np.float32)
    return Workload("plain_small", shape, params, ids, table, np.asarray([[16, 32, 32]], np.int32),
                    jz[:max(steps, 1)].copy(), np.asarray([[8, 32, 32, 32]], np.int32), steps, freq=2.45e9)


def plain_256(steps=2000, seed=12347, n=256):
    """n^3 (256 by default) with 8 nested seeded ellipsoids and a Jz sheet on x = 10."""
    shape = (n, n, n)
    params, jz = derive(shape, steps, 9.0e8, 2.0e-3, 0.99)
    rng = SplitMix64(seed)
    s = n / 256.0
    shapes = []
    for _ in range(8):
        a = [rng.uniform(20.0, 110.0) * s for _ in range(3)]
        c = [n / 2 + rng.uniform(-10.0, 10.0) * s for _ in range(3)]
        shapes.append((a[0] * a[1] * a[2], c, a))
    shapes.sort(key=lambda t: -t[0])  # largest first, so later ids nest inside
    table = np.concatenate([np.asarray([(0.0, 1.0, 1.0, 0.0)], np.float32), debye_table(rng, 8)])
    ids = np.zeros(shape[::-1], np.uint8)
    for m, (_, c, a) in enumerate(shapes, start=1):
        _paint_ellipsoid(ids, c, a, m)
    x = max(2, int(10 * s))
    jj, kk = np.meshgrid(np.arange(2, n - 2), np.arange(2, n - 2), indexing="ij")
    src = np.stack([np.full(jj.size, x), jj.ravel(), kk.ravel()], 1).astype(np.int32)
    h = n // 2
    probes = np.asarray([[8, h, h, h], [8, n // 4, h, h], [8, 3 * n // 4, h, h]], np.int32)
    return Workload("plain_256", shape, params, ids, table, src, jz[:max(steps, 1)].copy(), probes, steps,
                    slice_k=h, slice_every=100, seed=seed, freq=9.0e8)


def large_phantom(steps=5000, seed=12348, shape=(320, 320, 880)):
    """320x320x880 voxel phantom of 50 tissue ids (ellipsoids and z-cylinders)."""
    nx, ny, nz = shape
    params, jz = derive(shape, steps, 9.0e8, 2.0e-3, 0.99)
    rng = SplitMix64(seed)
    table = np.concatenate([np.asarray([(0.0, 1.0, 1.0, 0.0)], np.float32), debye_table(rng, 50)])
    ids = np.zeros((nz, ny, nx), np.uint8)
    # body: a large ellipsoid, then 49 seeded organs/limbs inside its box
    body_c = (nx / 2, ny / 2, nz / 2)
    body_a = (0.40 * nx, 0.30 * ny, 0.45 * nz)
    _paint_ellipsoid(ids, body_c, body_a, 1)
    for m in range(2, 51):
        cx = body_c[0] + rng.uniform(-0.7, 0.7) * body_a[0]
        cy = body_c[1] + rng.uniform(-0.7, 0.7) * body_a[1]
        cz = body_c[2] + rng.uniform(-0.8, 0.8) * body_a[2]
        if m % 2 == 0:
            a = (rng.uniform(8, 45), rng.uniform(8, 40), rng.uniform(10, 90))
            _paint_ellipsoid(ids, (cx, cy, cz), a, m)
        else:
            r = rng.uniform(5, 25)
            half = rng.uniform(20, 150)
            _paint_zcylinder(ids, (cx, cy), r, int(cz - half), int(cz + half), m)
    # dipole: 10 cells of Jz along z just outside the body surface on -x
    xs = max(2, int(body_c[0] - body_a[0]) - 4)
    zc = nz // 2
    src = np.asarray([[xs, ny // 2, z] for z in range(zc - 5, zc + 5)], np.int32)
    probes = np.asarray([[8, nx // 2, ny // 2, nz // 2], [8, xs + 8, ny // 2, zc], [6, nx // 2, ny // 2, nz // 4]],
                        np.int32)
    return Workload("large_phantom", shape, params, ids, table, src, jz[:max(steps, 1)].copy(), probes, steps,
                    slice_k=nz // 2, slice_every=500, seed=seed, freq=9.0e8)


# Vacuum cells kept between a phantom and every face.  First-order Mur with
# the reference's rp (from the default material, i.e. vacuum) diverges when
# a dielectric reaches the boundary: a Debye ellipsoid touching the faces
# grows by 10^29 within 1600 steps, one kept 10 cells inside stays bounded
# (the same holds with sigma = 0 or tau = 0, and without Mur it decays).
PHANTOM_MARGIN = 10


def ellipsoid_phantom(rng, shape, nmat):
    """Vacuum background plus nmat - 1 seeded ellipsoids painted largest
    first, inside the box PHANTOM_MARGIN cells in from every face (semi-axes
    15-45 % of the box edge, centres in the middle 40 %, each ellipsoid
    clipped to the box)."""
    nx, ny, nz = shape
    M = PHANTOM_MARGIN
    table = np.concatenate([np.asarray([(0.0, 1.0, 1.0, 0.0)], np.float32), debye_table(rng, nmat - 1)])
    shapes = []
    for _ in range(nmat - 1):
        box = [n - 1 - 2 * M for n in shape]
        a = [rng.uniform(0.15, 0.45) * L for L in box]
        c = [M + rng.uniform(0.3, 0.7) * L for L in box]
        a = [min(a[d], c[d] - M, shape[d] - 1 - M - c[d]) for d in range(3)]
        if min(a) < 1.0:
            continue
        shapes.append((a[0] * a[1] * a[2], c, a))
    shapes.sort(key=lambda t: -t[0])
    ids = np.zeros((nz, ny, nx), np.uint8)
    for m, (_, c, a) in enumerate(shapes, start=1):
        _paint_ellipsoid(ids, c, a, m)
    return ids, table


def _grid(name, rng, shape, steps, freq, nmat, dipole, seed):
    nx, ny, nz = shape
    params, jz = derive(shape, steps, freq, 2.0e-3, 0.99)
    ids, table = ellipsoid_phantom(rng, shape, nmat)
    x, y = rng.randint(2, nx - 3), rng.randint(2, ny - 3)
    if dipole and nz >= 14:
        z0 = rng.randint(1, nz - 12)
        src = np.asarray([[x, y, z] for z in range(z0, z0 + 10)], np.int32)
    else:
        src = np.asarray([[x, y, rng.randint(1, nz - 3)]], np.int32)
    probes = np.asarray([[8, nx // 2, ny // 2, nz // 2], [8, min(x + 3, nx - 1), y, int(src[0, 2])]], np.int32)
    return Workload(name, shape, params, ids, table, src, jz[:max(steps, 1)].copy(), probes, steps, seed=seed,
                    freq=freq)


def batched_set(count=64, seed=12349):
    """count uneven grids between 48^3 and 160x160x224 (see module doc)."""
    rng = SplitMix64(seed)
    out = []
    for q in range(count):
        shape = (rng.randint(48, 160), rng.randint(48, 160), rng.randint(48, 224))
        nmat = rng.randint(4, 12)
        freq = rng.uniform(3.0e8, 6.0e9)
        steps = rng.randint(500, 3000)
        out.append(_grid(f"batched_{q}", rng, shape, steps, freq, nmat, False, seed))
    return out


def box_phantom(shape, steps=100, seed=12351):
    """Vacuum plus four seeded Debye boxes (see module doc); painted by
    slicing, so a grid of 10^9 cells builds in seconds."""
    nx, ny, nz = shape
    params, jz = derive(shape, steps, 9.0e8, 2.0e-3, 0.99)
    rng = SplitMix64(seed)
    table = np.concatenate([np.asarray([(0.0, 1.0, 1.0, 0.0)], np.float32), debye_table(rng, 4)])
    ids = np.zeros((nz, ny, nx), np.uint8)
    M = PHANTOM_MARGIN
    for m in range(1, 5):
        lo = [M + int(rng.uniform(0.0, 0.4) * (n - 2 * M)) for n in (nx, ny, nz)]
        hi = [min(n - M, l + max(1, int(rng.uniform(0.1, 0.5) * (n - 2 * M)))) for l, n in zip(lo, (nx, ny, nz))]
        ids[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = m
    c = (nx // 2, ny // 2, nz // 2)
    src = np.asarray([c], np.int32)
    probes = np.asarray([[8, c[0], c[1], c[2]], [8, c[0] + 8, c[1], c[2]]], np.int32)
    return Workload("box_phantom", shape, params, ids, table, src, jz[:max(steps, 1)].copy(), probes, steps,
                    seed=seed, freq=9.0e8)


def served_stream(count=120, seed=12350, rate=40.0, f16_share=0.5):
    """[(arrival seconds, Workload, storage 'f32'|'f16')] of an open-loop
    Poisson stream (see module doc)."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for q in range(count):
        t += -np.log(1.0 - rng.uniform()) / rate
        shape = (rng.randint(32, 192), rng.randint(32, 192), rng.randint(32, 192))
        steps = rng.randint(200, 4000)
        storage = "f16" if rng.uniform() < f16_share else "f32"
        nmat = rng.randint(4, 12)
        freq = rng.uniform(3.0e8, 6.0e9)
        dipole = rng.uniform() < 0.5
        out.append((t, _grid(f"served_{q}", rng, shape, steps, freq, nmat, dipole, seed), storage))
    return out


def riken(ids, table95, steps=200, pad=16):
    """The RIKEN phantom (uint8 tissue ids (nz, ny, nx), id = tissue code)
    padded with `pad` vacuum cells on every side; see the module doc."""
    nz0, ny0, nx0 = ids.shape
    shape = (nx0 + 2 * pad, ny0 + 2 * pad, nz0 + 2 * pad)
    nx, ny, nz = shape
    params, jz = derive_reference(shape, steps)
    full = np.zeros((nz, ny, nx), np.uint8)
    full[pad:pad + nz0, pad:pad + ny0, pad:pad + nx0] = ids
    table = np.asarray(table95, np.float32).copy()
    table[0] = (0.0, 1.0, 1.0, 0.0)   # code 0 is air
    c = (nx // 2, ny // 2, nz // 2)
    src = np.asarray([c], np.int32)
    probes = np.asarray([[8, c[0], c[1], c[2]], [8, c[0], c[1] - 40, c[2]], [6, c[0], c[1], c[2] + 200]], np.int32)
    return Workload("riken", shape, params, full, table, src, jz[:max(steps, 1)].copy(), probes, steps, freq=3.0e9)


def derive_reference(shape, steps):
    """Parameters of data/input_params (3 GHz, pulse width 0.5, modulation 0,
    100 cells per wavelength) on another grid, at Courant 0.99."""
    L = lib()
    nx, ny, nz = shape
    text = params_text(shape, steps, 3.0e9, 0.0, 0.99, pulse_width=0.5, pmf=0.0,
                       sources=((nx // 2, ny // 2, nz // 2),))
    text = text.replace("cell_size: 0.000000000e+00\n", "")
    with tempfile.NamedTemporaryFile("w", suffix=".params", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        pp = L.initParamsWithPath(path.encode())
        L.setupSources(pp)
        p = pp.contents
        jz = np.ctypeslib.as_array(p.jz, shape=(p.jzCount,)).astype(np.float32).copy()
        out = FdtdParams()
        C.memmove(C.byref(out), C.byref(p), C.sizeof(FdtdParams))
        out.inputPath = out.outputPath = None
        out.sources = None
        out.jz = None
        out.sourcesCount = out.jzCount = 0
        L.deallocParams(pp)
    finally:
        os.unlink(path)
    return out, jz
