"""Pins of the oracle's set-up functions: parameter derivation, source
waveform, Mur permittivity planes, tissue table and phantom slices.

Anchors are closed forms evaluated independently in numpy and the known
answers derived for the reference's own data/input_params fixture
(fine 23, temp 3983, istart -3979, 347 samples, jz support [0,346],
peak |jz| 3.43, dx 1 mm, dt 1.9245e-12 s)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PARAMS = os.path.join(GOLDEN, "input_params")
SPECS = os.path.join(GOLDEN, "mat_specs_riken")
REF_PGM = "/root/reference/data/pgmdata/v1_00001.pgm"


def test_params_fixture_closed_forms():
    p = O.parse_params(PARAMS)
    assert (p.nx, p.ny, p.nz) == (100, 100, 100)
    assert p.iterations == 3
    assert p.epw == 100 and p.nsrc == 1
    assert list(p.src[:3]) == [49, 49, 49]  # 1-based 50 50 50 in the file
    assert (p.sigma, p.eps_s, p.eps_i, p.tau_d) == (0.0, 1.0, 1.0, 0.0)
    assert p.input_path == b"data/pgmdata" and p.output_path == b"output"
    c = 3.0e8
    lam = c / 3.0e9
    dx = lam / 100
    assert p.dx == pytest.approx(dx, rel=1e-7) and p.dy == p.dx and p.dz == p.dx
    assert p.dt == pytest.approx(dx / (c * np.sqrt(3.0)), rel=1e-7)
    assert p.dt == pytest.approx(1.9245e-12, rel=1e-4)
    assert p.mu0 == pytest.approx(4e-7 * np.pi, rel=1e-7)
    assert p.eps0 == pytest.approx(1.0 / (4e-7 * np.pi * c * c), rel=1e-6)


@pytest.mark.parametrize("tmax,expect", [(1, 3), (3, 3), (4, 6), (500, 501), (2000, 2001)])
def test_params_tmax_rounded_to_multiple_of_three(tmp_path, tmax, expect):
    text = open(PARAMS).read().replace("t_max: 3", f"t_max: {tmax}")
    f = tmp_path / "p"
    f.write_text(text)
    assert O.parse_params(str(f)).iterations == expect


def test_jz_known_answers_input_params():
    p = O.parse_params(PARAMS)
    jz, info = O.setup_jz(p.pulse_width, p.wave_frequency, p.pulse_modulation_frequency, p.dt, p.dx, p.dy, p.dz)
    assert info == dict(fine=23, temp=3983, istart=-3979, nraw=347)
    nz = np.nonzero(jz)[0]
    assert nz.min() == 0 and nz.max() == 346
    assert np.abs(jz).max() == pytest.approx(3.43, abs=0.01)


def _analytic_jz(pw, f, pmf, dt, dx, dy, dz, info, n):
    """Closed form of the waveform: g(t) = exp(-(t/T)^2) cos(2 pi pmf f t)
    sampled at t_m = (istart + m fine) dt/fine, scaled 1e-15/dt/3, then the
    central difference jz[n] ~ g'(t_{n+1}) * 1e-15/dt/3 * dt * dz/(dx dy dz)."""
    fine, temp, istart = info["fine"], info["temp"], info["istart"]
    sub = dt / fine
    T = (temp + 1.0) / 4.0 * sub
    w = 2 * np.pi * pmf * f
    t = (istart + (np.arange(n) + 1) * fine) * sub
    g = np.exp(-(t / T) ** 2)
    dg = g * (-2 * t / T ** 2 * np.cos(w * t) - w * np.sin(w * t))
    return dg * 1e-15 / dt / 3.0 * dt * dz / (dx * dy * dz)


@pytest.mark.parametrize("freq,cell,pmf", [(3.0e9, None, 0.0), (2.45e9, 2e-3, 1.0), (9e8, 2e-3, 1.0)])
def test_jz_matches_derivative_of_gaussian_pulse(freq, cell, pmf):
    p = O.parse_params(PARAMS)
    p.wave_frequency = freq
    p.pulse_modulation_frequency = pmf
    O.derive(p, 1.0 if cell is None else 0.99)
    if cell is not None:
        p.dx = p.dy = p.dz = cell
        p.dt = 0.99 * cell / (p.c * np.sqrt(3.0))
    jz, info = O.setup_jz(p.pulse_width, p.wave_frequency, p.pulse_modulation_frequency, p.dt, p.dx, p.dy, p.dz)
    n = info["nraw"] - 2
    ref = _analytic_jz(p.pulse_width, p.wave_frequency, p.pulse_modulation_frequency, p.dt, p.dx, p.dy, p.dz, info, n)
    scale = np.abs(ref).max()
    # central differences of float samples: O(dt^2) truncation + float rounding
    assert np.abs(jz[:n] - ref).max() < 2e-3 * scale
    # a central difference telescopes: the sum is the end samples' difference
    assert abs(float(np.sum(jz[: info["nraw"] + 2], dtype=np.float64))) < 1e-4 * np.abs(jz).sum()


def test_rp_matches_debye_permittivity():
    rng = np.random.default_rng(5)
    nx, ny, nz = 7, 6, 5
    f = 9.0e8
    pi = np.float32(np.pi)
    eps0 = np.float32(8.841941e-12)
    sg = rng.uniform(0.0, 2.5, (nz, ny, nx)).astype(np.float32)
    ei = rng.uniform(1, 45, (nz, ny, nx)).astype(np.float32)
    es = (ei + rng.uniform(0, 40, (nz, ny, nx))).astype(np.float32)
    tau = rng.uniform(0, 5e-10, (nz, ny, nx)).astype(np.float32)
    rps = O.setup_rp((nx, ny, nz), f, pi, eps0, sg, es, ei, tau)
    w = 2 * np.pi * np.float64(f)
    d = lambda a: a.astype(np.float64)
    # x planes: c1 and c2 real (no I): ei + (es-ei)/(1 + w tau) - sigma/(w eps0)
    rx = d(ei) + (d(es) - d(ei)) / (1 + w * d(tau)) - d(sg) / (w * float(eps0))
    # y/z planes: Re(ei + (es-ei)/(1 + j w tau) - j sigma/(w eps0))
    ryz = d(ei) + (d(es) - d(ei)) / (1 + (w * d(tau)) ** 2)
    rpx0, rpy0, rpz0, rpxE, rpyE, rpzE = [r.astype(np.float64) for r in rps]
    close = lambda a, b: np.testing.assert_allclose(a, b, rtol=2e-6, atol=2e-6 * np.abs(b).max())
    close(rpx0.reshape(nz, ny, 2), rx[:, :, 0:2])
    close(rpxE.reshape(nz, ny, 2), rx[:, :, nx - 2:])
    close(rpy0.reshape(nz, 2, nx), ryz[:, 0:2, :])
    close(rpyE.reshape(nz, 2, nx), ryz[:, ny - 2:, :])
    close(rpz0.reshape(2, ny, nx), ryz[0:2])
    close(rpzE.reshape(2, ny, nx), ryz[nz - 2:])


def test_rp_vacuum_is_one():
    shape = (5, 4, 3)
    ones = np.ones(shape[::-1], np.float32)
    zeros = np.zeros(shape[::-1], np.float32)
    for r in O.setup_rp(shape, 3e9, np.float32(np.pi), np.float32(8.841941e-12), zeros, ones, ones, zeros):
        assert np.all(r == 1.0)


def test_specs_fixture():
    specs, rows = O.load_specs(SPECS)
    assert rows == 53
    np.testing.assert_array_equal(specs[0], np.float32([0, 1, 1, 0]))
    np.testing.assert_array_equal(specs[1], np.float32([0.59515458, 56.44398499, 33.05709076, 3.51972652e-11]))
    np.testing.assert_array_equal(specs[94], np.float32([0.80907071, 60.54281044, 27.70847893, 1.77068464e-11]))
    np.testing.assert_array_equal(specs[17], 0)  # not in the table
    listed = [int(l.split()[0]) for l in open(SPECS) if l[:1].isdigit()]
    assert len(listed) == 53 and all(specs[c].any() for c in listed if c not in (0, 37, 82, 87, 88))


def _write_pgm(path, w, h, values, comment="# test"):
    with open(path, "w") as f:
        f.write(f"P2\n{comment}\n{w} {h}\n94\n")
        f.write("\n".join(str(int(v)) for v in values) + "\n")


def test_pgm_reads_first_nxny_codes_ignoring_width(tmp_path):
    rng = np.random.default_rng(2)
    vals = rng.integers(0, 95, 12 * 9)
    p = tmp_path / "v1_00001.pgm"
    _write_pgm(p, 12, 9, vals)
    codes, w, h = O.load_pgm_codes(str(p), 5 * 4)  # nx=5, ny=4: the first 20 values
    assert (w, h) == (12, 9)
    np.testing.assert_array_equal(codes, vals[:20])


@pytest.mark.skipif(not os.path.exists(REF_PGM), reason="reference phantom slices not present on this host")
def test_pgm_reference_slice_header():
    codes, w, h = O.load_pgm_codes(REF_PGM, 100 * 100)
    assert (w, h) == (265, 490)
    assert codes.min() >= 0 and codes.max() <= 94
