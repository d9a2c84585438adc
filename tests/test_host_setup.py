"""The library's host services (include/fdtd.h) against the oracle, on the
CPU: parameter parsing and derivation, the source waveform, the Mur
permittivity planes, the tissue table + phantom slices, and the text result
formats.  Bit equality for every number; exact text for the writers."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fdtd_amd as F
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PARAMS = os.path.join(GOLDEN, "input_params")
SPECS = os.path.join(GOLDEN, "mat_specs_riken")
libc = C.CDLL(None)
libc.malloc.restype = C.c_void_p
libc.malloc.argtypes = [C.c_size_t]


def _product_params(path):
    pp = F.lib().initParamsWithPath(path.encode())
    return pp


def _f32bits(x):
    return np.float32(x).view(np.uint32)


def test_params_match_oracle_bitwise():
    pp = _product_params(PARAMS)
    p = pp.contents
    o = O.parse_params(PARAMS)
    assert (p.nx, p.ny, p.nz, p.iterationsCount, p.elementsPerWave, p.sourcesCount) == \
        (o.nx, o.ny, o.nz, o.iterations, o.epw, o.nsrc)
    assert [p.sources[q] for q in range(3)] == list(o.src[:3])
    for a, b in [("waveFrequency", "wave_frequency"), ("pulseWidth", "pulse_width"),
                 ("pulseModulationFrequency", "pulse_modulation_frequency"), ("defaultSigma", "sigma"),
                 ("defaultEpsS", "eps_s"), ("defaultEpsI", "eps_i"), ("defaultTauD", "tau_d"), ("pi", "pi"),
                 ("c", "c"), ("timeskip", "timeskip"), ("lambda_", "lambda_"), ("dt", "dt"), ("dx", "dx"),
                 ("dy", "dy"), ("dz", "dz"), ("mu0", "mu0"), ("eps0", "eps0")]:
        assert _f32bits(getattr(p, a)) == _f32bits(getattr(o, b)), a
    assert p.inputPath == o.input_path and p.outputPath == o.output_path
    F.lib().deallocParams(pp)


def test_params_extension_keys(tmp_path):
    """courant_factor and cell_size after ftau_d (the reference stops
    reading there): dt = S dx / (c sqrt 3)."""
    text = open(PARAMS).read().rstrip("\n") + "\ncourant_factor: 0.99\ncell_size: 2.0e-3\n"
    f = tmp_path / "p"
    f.write_text(text)
    pp = _product_params(str(f))
    p = pp.contents
    assert p.dx == np.float32(2e-3) and p.dy == p.dx and p.dz == p.dx
    assert p.timeskip == np.float32(0.99)
    assert p.dt == pytest.approx(0.99 * 2e-3 / (3e8 * np.sqrt(3)), rel=1e-6)
    # the oracle ignores the extension lines: the reference's own fields still agree
    o = O.parse_params(str(f))
    assert (p.nx, p.iterationsCount, p.sourcesCount) == (o.nx, o.iterations, o.nsrc)
    F.lib().deallocParams(pp)


def test_params_missing_file_exits():
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path.insert(0, %r); import fdtd_amd as F; "
            "F.lib().initParamsWithPath(b'/nonexistent/params')" % root)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=root)
    assert r.returncode != 0
    assert "Couldn't open file /nonexistent/params" in r.stdout


@pytest.mark.parametrize("freq,pmf,cell,courant", [(3.0e9, 0.0, None, None), (2.45e9, 1.0, 2e-3, 0.99),
                                                   (9.0e8, 1.0, 2e-3, 0.99), (6.0e9, 1.0, 1e-3, 0.9)])
def test_jz_matches_oracle_bitwise(tmp_path, freq, pmf, cell, courant):
    text = open(PARAMS).read()
    text = text.replace("wave_frequency: 3.0e+9", f"wave_frequency: {freq:.6e}")
    text = text.replace("pulse_modulation_frequency: 0", f"pulse_modulation_frequency: {pmf}")
    if cell is not None:
        text = text.rstrip("\n") + f"\ncourant_factor: {courant}\ncell_size: {cell}\n"
    f = tmp_path / "p"
    f.write_text(text)
    pp = _product_params(str(f))
    F.lib().setupSources(pp)
    p = pp.contents
    got = np.ctypeslib.as_array(p.jz, shape=(p.jzCount,)).copy()
    ref, info = O.setup_jz(p.pulseWidth, p.waveFrequency, p.pulseModulationFrequency, p.dt, p.dx, p.dy, p.dz)
    assert p.jzCount == 1 << 16
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.count_nonzero(got) > 10
    F.lib().deallocParams(pp)


def _field_struct(shape, mats):
    nx, ny, nz = shape
    keep = []
    fld = F.FdtdField()
    for name, arr in zip(("sigma", "epsS", "epsI", "tauD"), mats):
        a = np.ascontiguousarray(arr, np.float32).ravel().copy()
        keep.append(a)
        setattr(fld, name, a.ctypes.data_as(C.POINTER(C.c_float)))
    rps = [np.zeros(n, np.float32) for n in (2 * ny * nz, nx * 2 * nz, nx * ny * 2) * 2]
    for name, a in zip(("rpx0", "rpy0", "rpz0", "rpxEnd", "rpyEnd", "rpzEnd"), rps):
        setattr(fld, name, a.ctypes.data_as(C.POINTER(C.c_float)))
    return fld, keep, rps


def test_rp_planes_match_oracle_bitwise():
    rng = np.random.default_rng(11)
    shape = (9, 7, 6)
    nx, ny, nz = shape
    sg = rng.uniform(0, 2.5, shape[::-1]).astype(np.float32)
    ei = rng.uniform(1, 45, shape[::-1]).astype(np.float32)
    es = (ei + rng.uniform(0, 35, shape[::-1])).astype(np.float32)
    tau = rng.uniform(0, 8e-10, shape[::-1]).astype(np.float32)
    pp = _product_params(PARAMS)
    p = pp.contents
    p.nx, p.ny, p.nz = nx, ny, nz
    p.waveFrequency = np.float32(9e8)
    fld, keep, rps = _field_struct(shape, (sg, es, ei, tau))
    F.lib().setupMurBoundary(pp, C.byref(fld))
    ref = O.setup_rp(shape, p.waveFrequency, p.pi, p.eps0, sg, es, ei, tau)
    for got, want in zip(rps, ref):
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    F.lib().deallocParams(pp)


def _write_pgm(path, w, h, values):
    with open(path, "w") as f:
        f.write(f"P2\n# test\n{w} {h}\n94\n")
        f.write("\n".join(str(int(v)) for v in values) + "\n")


def test_load_materials_matches_oracle(tmp_path):
    """Tissue table + nz PGM slices (width 13, wider than nx: the
    reference's width-ignoring read order) into the four per-voxel arrays."""
    rng = np.random.default_rng(12)
    shape = (6, 5, 4)
    nx, ny, nz = shape
    specs, _ = O.load_specs(SPECS)
    listed = [int(l.split()[0]) for l in open(SPECS) if l[:1].isdigit()]
    slices = []
    for k in range(nz):
        vals = rng.choice(listed, 13 * 9)
        _write_pgm(tmp_path / f"v1_{k + 1:05d}.pgm", 13, 9, vals)
        slices.append(O.load_pgm_codes(str(tmp_path / f"v1_{k + 1:05d}.pgm"), nx * ny)[0])
    codes = np.stack(slices).reshape(nz, ny, nx)
    pp = _product_params(PARAMS)
    p = pp.contents
    p.nx, p.ny, p.nz = nx, ny, nz
    fld, keep, _ = _field_struct(shape, [np.zeros(shape[::-1])] * 4)
    F.lib().loadMaterials(pp, C.byref(fld), SPECS.encode(), str(tmp_path).encode())
    got = dict(zip(("sigma", "epsS", "epsI", "tauD"), keep))
    for col, name in enumerate(("sigma", "epsS", "epsI", "tauD")):
        want = specs[codes, col].astype(np.float32).ravel()
        np.testing.assert_array_equal(got[name], want)
    F.lib().deallocParams(pp)


def _malloc_copy(a):
    a = np.ascontiguousarray(a, np.float32)
    ptr = libc.malloc(a.nbytes)
    C.memmove(ptr, a.ctypes.data, a.nbytes)
    return ptr


def test_writers_formats(tmp_path):
    rng = np.random.default_rng(13)
    shape = (5, 4, 3)
    nx, ny, nz = shape
    fields = [(rng.standard_normal(shape[::-1]) * 10.0 ** rng.integers(-12, 3)).astype(np.float32) for _ in range(9)]
    pp = _product_params(PARAMS)
    p = pp.contents
    p.nx, p.ny, p.nz = nx, ny, nz
    out = str(tmp_path).encode()
    keep_out = C.create_string_buffer(out + b"\0" * 8)
    p.outputPath = C.cast(keep_out, C.c_char_p)
    src = (C.c_int * 3)(2, 1, 1)
    p.sourcesCount = 1
    p.sources = C.cast(src, C.POINTER(C.c_int))
    order = [3, 4, 5, 0, 1, 2, 6, 7, 8]  # argument order: H, D, E
    F.lib().writeResults3d(pp, *[_malloc_copy(fields[q]) for q in order], 7)
    names = sorted(os.listdir(tmp_path))
    assert len(names) == 9 * nz
    assert "H_field_x_00007_z00000.out" in names and "E_field_z_00007_z00002.out" in names
    txt = open(tmp_path / "D_field_y_00007_z00001.out").read()
    lines = txt.split("\n")[:-1]
    assert len(lines) == ny
    assert lines[0] == "".join(" % 9.3E" % v for v in fields[1][1, 0])
    # line probes through the source: " %3d %3d %3d" then D, H, E columns
    F.lib().writeResults(pp, *[_malloc_copy(fields[q]) for q in order], 7)
    xs = open(tmp_path / "E_field_x_00008.out").read().split("\n")[:-1]
    assert len(xs) == nx
    i, j, k = 3, 1, 1
    want = " %3d %3d %3d" % (i + 1, j + 1, k + 1) + "".join(" % 9.3E" % fields[q][k, j, i] for q in range(9))
    assert xs[i] == want
    assert re.fullmatch(r"(\s+-?\d\.\d{3}E[+-]\d{2})+", " " + xs[0][13:])
    p.outputPath = None
    p.sources = None
    p.sourcesCount = 0
    F.lib().deallocParams(pp)
