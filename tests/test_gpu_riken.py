"""The reference's own phantom on the GPU.  Six slices of data/pgmdata
(cropped to 100x100 by tests/golden/make_riken_crop.py; 12 tissue codes) are
read in the reference's order (fdtd.cu:875-939: nx*ny codes per slice, the
tissue table's rows by code) and stepped two ways:
  * fdtd_load_phantom_ids + fdtd_load_tissue_table into a fused context;
  * the `fdtd` CLI (parameter file, tissue table, slices in; text out), fused
    and --legacy.
Both must equal the oracle's FP32 transcription fed the same slices bit for
bit: every field of the newest level, every probe sample, and every
character of the written slice files."""
import os

import numpy as np
import pytest

import fdtd_amd as F
from oracle import oracle as O
from tests.helpers import bits_equal, first_diff
from tests.test_gpu_cli import FIELD_FILES, _run, _slice_text

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CROP = os.path.join(GOLDEN, "riken_crop")
SPECS = os.path.join(GOLDEN, "mat_specs_riken")
SHAPE = (100, 100, 6)
SOURCE = (50, 40, 4)  # 1-based, as in the parameter file
STEPS = 60


def _params_text(tmp_path):
    text = open(os.path.join(GOLDEN, "input_params")).read()
    text = text.replace("nx_ny_nz: 100 100 100", "nx_ny_nz: %d %d %d" % SHAPE)
    text = text.replace("t_max: 3", f"t_max: {STEPS}")
    text = text.replace("env_set_dir: data/pgmdata", f"env_set_dir: {CROP}")
    text = text.replace("source_location: 50 50 50", "source_location: %d %d %d" % SOURCE)
    return tmp_path / "params", text


def _oracle(params_path, steps, probes=()):
    """The reference's pipeline on the oracle: parameters and waveform from
    the file, rp from the default material, the crop's codes through the
    tissue table, jz[3 floor(n/3)] from the first 512 samples, FIXED Debye.
    Returns the nine newest fields after each step and the probe series."""
    o = O.parse_params(str(params_path))
    nx, ny, nz = SHAPE
    assert (o.nx, o.ny, o.nz) == SHAPE
    full = lambda v: np.full((nz, ny, nx), v, np.float32)
    rps = O.setup_rp(SHAPE, o.wave_frequency, o.pi, o.eps0, full(o.sigma), full(o.eps_s), full(o.eps_i),
                     full(o.tau_d))
    specs, _ = O.load_specs(SPECS)
    codes = np.stack([O.load_pgm_codes(os.path.join(CROP, f"v1_{k + 1:05d}.pgm"), nx * ny)[0]
                      for k in range(nz)]).reshape(nz, ny, nx)
    mats = specs[codes]
    jz, _ = O.setup_jz(o.pulse_width, o.wave_frequency, o.pulse_modulation_frequency, o.dt, o.dx, o.dy, o.dz)
    sim = O.Sim(SHAPE, o.dt, o.dx, o.dy, o.dz, o.mu0, o.eps0, prec=32, debye=O.DEBYE_FIXED)
    sim.set_materials(mats[..., 0], mats[..., 1], mats[..., 2], mats[..., 3])
    sim.set_rp(rps)
    src = np.asarray(o.src[:3 * o.nsrc], np.int32).reshape(-1, 3)
    sim.set_sources(src, jz[:512], jz_mode=0)
    states, series = [], []
    for _ in range(steps):
        series.append(sim.run(1, probes))
        states.append([sim.get(f) for f in range(9)])
    return states, np.concatenate(series), codes, o


def test_riken_crop_context_bit_equal(tmp_path):
    """Slices -> uint8 tissue ids -> fused context, against the oracle."""
    params_path, text = _params_text(tmp_path)
    params_path.write_text(text)
    nx, ny, nz = SHAPE
    probes = np.asarray([[8, 49, 39, 3], [6, 55, 42, 2], [2, 49, 39, 3], [4, 45, 35, 1], [7, 99, 39, 3],
                         [8, 49, 0, 3]], np.int32)
    states, series, codes, o = _oracle(params_path, STEPS, probes)
    ids = F.load_phantom_ids(CROP, SHAPE, F.PGM_REFERENCE, threads=4)
    np.testing.assert_array_equal(ids, codes)
    table, rows = F.load_tissue_table(SPECS)
    assert rows == 53
    L = F.lib()
    pp = L.initParamsWithPath(str(params_path).encode())
    L.setupSources(pp)
    p = pp.contents
    jz = np.ctypeslib.as_array(p.jz, shape=(p.jzCount,)).astype(np.float32).copy()
    src = np.ctypeslib.as_array(p.sources, shape=(3 * p.sourcesCount,)).reshape(-1, 3).copy()
    params = F.make_params(SHAPE, p.dt, p.dx, p.dy, p.dz, mu0=p.mu0, eps0=p.eps0)
    L.deallocParams(pp)
    print(f"\nriken crop {SHAPE}, {STEPS} steps, source {src.tolist()}", flush=True)
    with F.Context(params, jz_mode=F.JZ_HOLD3, jz_limit=512) as ctx:
        ctx.set_materials(ids, table)
        ctx.set_sources(src, jz)
        ctx.set_probes(probes, STEPS)
        ctx.run(STEPS)
        got = ctx.probes()
        assert np.array_equal(got.view(np.uint64), series.view(np.uint64)), "probe series differ"
        for f in range(9):
            a = ctx.get(f)
            assert bits_equal(a, states[-1][f]), f"{F.FIELDS[f]}: {first_diff(a, states[-1][f])}"
    # the pulse reached tissue of several kinds
    ez = states[-1][8]
    assert len(np.unique(codes[np.abs(ez) > 0])) >= 3


def test_riken_crop_cli_matches_oracle(tmp_path):
    """The CLI on the crop, fused and --legacy: the slice files of steps 20,
    40 and 60 equal the oracle's, character for character."""
    params_path, text = _params_text(tmp_path)
    print(f"\ncli: riken crop {SHAPE}, {STEPS} steps, slices every 20", flush=True)
    fused = _run(params_path, text, tmp_path / "out_fused", "--every", "20")
    legacy = _run(params_path, text, tmp_path / "out_legacy", "--every", "20", "--legacy")
    nz = SHAPE[2]
    assert len(fused) == 3 * 9 * nz
    assert fused == legacy
    states, _, _, _ = _oracle(params_path, STEPS)
    for it in (19, 39, 59):
        for f, (kind, comp) in enumerate(FIELD_FILES):
            for k in range(nz):
                name = f"{kind}_field_{comp}_{it:05d}_z{k:05d}.out"
                assert fused[name] == _slice_text(states[it][f][k]), name
