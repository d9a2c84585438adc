"""The `fdtd` command-line driver end to end (fdtd.cu:46-544): parameter
file, tissue table and PGM slices in, text result files out.  A 24x20x16
grid with a random tissue phantom runs 6 steps; the slice files (and the
line files of --lines) must equal, character for character, the same files
formatted from the oracle's FP32 transcription fed the same inputs, and the
legacy path (the five drop-in entry points) must write the same bytes as the
fused path."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fdtd_amd", "fdtd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
SPECS = os.path.join(GOLDEN, "mat_specs_riken")
SHAPE = (24, 20, 16)
SOURCE = (12, 10, 8)  # 1-based, as in the parameter file
FIELD_FILES = [("D", "x"), ("D", "y"), ("D", "z"), ("H", "x"), ("H", "y"), ("H", "z"), ("E", "x"), ("E", "y"),
               ("E", "z")]


def _case(tmp_path, t_max=6):
    nx, ny, nz = SHAPE
    pgm = tmp_path / "pgm"
    pgm.mkdir()
    rng = np.random.default_rng(5)
    listed = [int(l.split()[0]) for l in open(SPECS) if l[:1].isdigit()]
    for k in range(nz):
        vals = rng.choice(listed, nx * ny)
        with open(pgm / f"v1_{k + 1:05d}.pgm", "w") as f:
            f.write(f"P2\n# test\n{nx} {ny}\n94\n" + "\n".join(str(int(v)) for v in vals) + "\n")
    text = open(os.path.join(GOLDEN, "input_params")).read()
    text = text.replace("nx_ny_nz: 100 100 100", f"nx_ny_nz: {nx} {ny} {nz}")
    text = text.replace("t_max: 3", f"t_max: {t_max}")
    text = text.replace("env_set_dir: data/pgmdata", f"env_set_dir: {pgm}")
    text = text.replace("source_location: 50 50 50", "source_location: %d %d %d" % SOURCE)
    params = tmp_path / "params"
    return params, text


def _run(params, text, out, *args):
    out.mkdir()
    params.write_text(text.replace("output_dir: output", f"output_dir: {out}"))
    assert os.path.exists(CLI), f"{CLI} is missing: build with __graft_entry__.build()"
    r = subprocess.run([CLI, str(params), "--specs", SPECS, "--quiet", *args], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return {n: open(out / n).read() for n in sorted(os.listdir(out))}


def _oracle(params_path, steps):
    """The reference's pipeline on the oracle: rp from the default material
    (before the phantom loads), tissue table + slices, jz[3 floor(n/3)]
    from the first 512 waveform samples, FIXED Debye."""
    o = O.parse_params(str(params_path))
    nx, ny, nz = SHAPE
    assert (o.nx, o.ny, o.nz) == SHAPE
    full = lambda v: np.full((nz, ny, nx), v, np.float32)
    rps = O.setup_rp(SHAPE, o.wave_frequency, o.pi, o.eps0, full(o.sigma), full(o.eps_s), full(o.eps_i),
                     full(o.tau_d))
    specs, _ = O.load_specs(SPECS)
    pgm = o.input_path.decode() if isinstance(o.input_path, bytes) else o.input_path
    codes = np.stack([O.load_pgm_codes(os.path.join(pgm, f"v1_{k + 1:05d}.pgm"), nx * ny)[0]
                      for k in range(nz)]).reshape(nz, ny, nx)
    mats = specs[codes]
    jz, _ = O.setup_jz(o.pulse_width, o.wave_frequency, o.pulse_modulation_frequency, o.dt, o.dx, o.dy, o.dz)
    sim = O.Sim(SHAPE, o.dt, o.dx, o.dy, o.dz, o.mu0, o.eps0, prec=32, debye=O.DEBYE_FIXED)
    sim.set_materials(mats[..., 0], mats[..., 1], mats[..., 2], mats[..., 3])
    sim.set_rp(rps)
    src = np.asarray(o.src[:3 * o.nsrc], np.int32).reshape(-1, 3)
    sim.set_sources(src, jz[:512], jz_mode=0)
    states = []
    for _ in range(steps):
        sim.step()
        states.append([sim.get(f) for f in range(9)])
    return states, src


def _slice_text(a):
    return "".join("".join(" % 9.3E" % v for v in row) + "\n" for row in a)


def test_cli_slices_match_oracle_and_legacy(tmp_path):
    print("cli: fused and --legacy, 24x20x16, 6 steps, slices every 3", flush=True)
    params, text = _case(tmp_path)
    fused = _run(params, text, tmp_path / "out_fused", "--every", "3")
    legacy = _run(params, text, tmp_path / "out_legacy", "--every", "3", "--legacy")
    nz = SHAPE[2]
    assert len(fused) == 2 * 9 * nz
    assert fused == legacy
    states, _ = _oracle(params, 6)
    for it in (2, 5):
        for f, (kind, comp) in enumerate(FIELD_FILES):
            for k in (0, 7, nz - 1):
                name = f"{kind}_field_{comp}_{it:05d}_z{k:05d}.out"
                assert fused[name] == _slice_text(states[it][f][k]), name
    assert any(v.strip("0E+. -\n") for v in fused.values()), "all slices are zero"


def test_cli_line_files_match_oracle(tmp_path):
    print("cli: --lines, 24x20x16, 3 steps", flush=True)
    params, text = _case(tmp_path, t_max=3)
    got = _run(params, text, tmp_path / "out", "--lines")
    states, src = _oracle(params, 3)
    nx, ny, nz = SHAPE
    i0, j0, k0 = (int(v) for v in src[0])
    for it in range(3):
        fields = states[it]
        # one line per cell along each axis through the source, 1-based coordinates
        for axis, n in (("x", nx), ("y", ny), ("z", nz)):
            lines = got[f"E_field_{axis}_{it + 1:05d}.out"].split("\n")[:-1]
            assert len(lines) == n
            for q in (0, n // 2, n - 1):
                i, j, k = (q, j0, k0) if axis == "x" else ((i0, q, k0) if axis == "y" else (i0, j0, q))
                want = " %3d %3d %3d" % (i + 1, j + 1, k + 1) + "".join(" % 9.3E" % fields[f][k, j, i]
                                                                       for f in range(9))
                assert lines[q] == want, (axis, it, q)


@pytest.mark.parametrize("legacy", [False, True])
def test_cli_overlapped_output_equals_sync(tmp_path, legacy):
    """Dumps written by the overlapped pipeline (device staging, pinned
    D2H on a side stream, copy and writer threads) equal, byte for byte,
    those written on the stepping thread after each step (--sync); with one
    writer in flight and with seven."""
    print(f"cli: overlapped vs --sync, every step, legacy={legacy}", flush=True)
    params, text = _case(tmp_path, t_max=9)
    extra = ("--legacy",) if legacy else ()
    sync = _run(params, text, tmp_path / "out_sync", "--every", "1", "--sync", *extra)
    one = _run(params, text, tmp_path / "out_one", "--every", "1", "--writers", "1", *extra)
    seven = _run(params, text, tmp_path / "out_seven", "--every", "1", *extra)
    assert len(sync) == 9 * 9 * SHAPE[2]
    assert one == sync and seven == sync
    lines = _run(params, text, tmp_path / "out_lines", "--every", "2", "--lines", *extra)
    lines_sync = _run(params, text, tmp_path / "out_lines_sync", "--every", "2", "--lines", "--sync", *extra)
    assert lines == lines_sync and len(lines) == 4 * 3
