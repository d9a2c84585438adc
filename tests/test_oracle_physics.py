"""Pins of the oracle's update equations by what the mathematics fixes:
discrete divergence identities of the Yee curls, the Debye recurrence's
closed-form growth factor and frequency response, the textbook form of
first-order Mur in vacuum, absorption against a closed box, source
linearity and the 3-step sample hold, mirror symmetry, and FP32 vs FP64
agreement.  None of these retypes the oracle's formulas."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import EPS0, MU0, rel_l2

DT = 3.8105118e-12
DX = 2.0e-3


def _grid(shape):
    nx, ny, nz = shape
    return O.grid(nx, ny, nz, DT, DX, DX, DX, MU0, EPS0)


def _z(shape):
    nx, ny, nz = shape
    return np.zeros((nz, ny, nx), np.float32)


def test_h_update_keeps_discrete_div_b_zero():
    """div(curl E) = 0 on the Yee grid: starting from H = 0, one H update
    from random E leaves the corner divergence zero up to rounding."""
    rng = np.random.default_rng(0)
    shape = (12, 11, 10)
    ex, ey, ez = [rng.uniform(-1, 1, shape[::-1]).astype(np.float32) for _ in range(3)]
    hx, hy, hz = _z(shape), _z(shape), _z(shape)
    O.lib().or_update_h_f32(C.byref(_grid(shape)), hx, hy, hz, ex, ey, ez)
    d = lambda a: a.astype(np.float64)
    # corner (i+1/2, j+1/2, k+1/2) for i in [1, nx-2), j in [1, ny-2), k in [1, nz-2)
    div = (d(hx)[1:-2, 1:-2, 2:-1] - d(hx)[1:-2, 1:-2, 1:-2]) + (d(hy)[1:-2, 2:-1, 1:-2] - d(hy)[1:-2, 1:-2, 1:-2]) \
        + (d(hz)[2:-1, 1:-2, 1:-2] - d(hz)[1:-2, 1:-2, 1:-2])
    scale = np.abs(d(hx)).max()
    assert scale > 0
    assert np.abs(div).max() < 1e-5 * scale


def test_d_update_keeps_div_d_zero_and_source_charge():
    """div(curl H) = 0 at nodes away from sources; at a source node the
    divergence is -jz (and +jz at the node above), the charge the soft
    current leaves behind."""
    rng = np.random.default_rng(1)
    shape = (12, 11, 10)
    hx, hy, hz = [rng.uniform(-1, 1, shape[::-1]).astype(np.float32) for _ in range(3)]
    dxs, dys, dzs = _z(shape), _z(shape), _z(shape)
    dxt, dyt, dzt = _z(shape), _z(shape), _z(shape)
    g = _grid(shape)
    O.lib().or_update_d_f32(C.byref(g), dxt, dyt, dzt, dxs, dys, dzs, hx, hy, hz)
    src = np.asarray([[5, 6, 4]], np.int32)
    jz = 0.7
    O.lib().or_update_sources_f32(C.byref(g), dzt, dzs, hx, hy, 1, src.ravel(), jz)
    d = lambda a: a.astype(np.float64)
    # node (i, j, k) for i, j, k in [1, n-1)
    div = (d(dxt)[1:-1, 1:-1, 1:-1] - d(dxt)[1:-1, 1:-1, :-2]) + (d(dyt)[1:-1, 1:-1, 1:-1] - d(dyt)[1:-1, :-2, 1:-1]) \
        + (d(dzt)[1:-1, 1:-1, 1:-1] - d(dzt)[:-2, 1:-1, 1:-1])
    scale = np.abs(d(dxt)).max()
    x, y, z = src[0]
    assert div[z - 1, y - 1, x - 1] == pytest.approx(-jz, rel=1e-5)
    assert div[z, y - 1, x - 1] == pytest.approx(jz, rel=1e-5)
    div[z - 1, y - 1, x - 1] = 0
    div[z, y - 1, x - 1] = 0
    assert np.abs(div).max() < 1e-5 * scale


def _e_point(sigma, eps_s, eps_i, tau, debye, drive, steps, prec=32):
    """Run the E update at one interior cell of a 3^3 grid with D forced to
    drive(n); returns the E series."""
    shape = (3, 3, 3)
    g = _grid(shape)
    mat = [np.full(shape[::-1], v, np.float32) for v in (sigma, eps_i, eps_s, tau)]  # sigma, epsI, epsS, tau
    E = [_z(shape) for _ in range(9)]     # 3 slots x 3 components
    D = [_z(shape) for _ in range(9)]
    out = []
    for n in range(steps):
        t, s1, s2 = n % 3, (n + 2) % 3, (n + 1) % 3
        for q in range(3):
            D[3 * t + q][1, 1, 1] = drive(n + 1)
        args = [E[3 * t + q] for q in range(3)] + [E[3 * s1 + q] for q in range(3)] + [E[3 * s2 + q] for q in range(3)]
        args += [D[3 * t + q] for q in range(3)] + [D[3 * s1 + q] for q in range(3)] + [D[3 * s2 + q] for q in range(3)]
        O.lib().or_update_e_f32(C.byref(g), debye, *args, *mat)
        out.append(float(E[3 * t + 0][1, 1, 1]))
    return np.asarray(out)


def _coeffs(sigma, eps_s, eps_i, tau, debye):
    a = 2 * EPS0 * eps_i * tau
    b = a + 2 * DT * (EPS0 * eps_s + sigma * tau)
    c = 1 / (b + sigma * DT * DT)
    d = (2 * b if debye == O.DEBYE_REF else a + b) - sigma * DT * DT
    return a, c, d


@pytest.mark.parametrize("mat,debye,expect", [
    ((0.0, 1.0, 1.0, 0.0), O.DEBYE_REF, 2.0),              # vacuum, the reference's d: doubling each step
    ((0.8, 55.0, 40.0, 8e-12), O.DEBYE_REF, 1.6274),        # the Debye sphere material
    ((0.0, 1.0, 1.0, 0.0), O.DEBYE_FIXED, 1.0),
    ((0.8, 55.0, 40.0, 8e-12), O.DEBYE_FIXED, 0.9938),
])
def test_e_recurrence_growth_factor(mat, debye, expect):
    """With D held constant the E update is E' = c(d E - a E_prev); its
    per-step growth is the dominant root of r^2 - c d r + c a = 0."""
    sigma, es, ei, tau = mat
    series = _e_point(sigma, es, ei, tau, debye, lambda n: 1e-10 if n <= 1 else 1e-10, 80, prec=32)
    a, c, d = _coeffs(sigma, es, ei, tau, debye)
    roots = np.roots([1.0, -c * d, c * a])
    root = max(roots, key=abs).real
    assert root == pytest.approx(expect, abs=5e-4)
    if abs(root) > 1.0001:
        ratio = series[-1] / series[-2]
        assert ratio == pytest.approx(root, rel=1e-3)


@pytest.mark.parametrize("mat", [(0.8, 55.0, 40.0, 8e-12), (0.3, 41.3, 24.4, 3.36e-11), (1.5, 70.0, 10.3, 8.3e-12)])
def test_fixed_debye_recovers_permittivity(mat):
    """Driven at omega, the FIXED recurrence's steady state satisfies
    D = eps0 eps(omega) E with eps = eps_inf + (eps_s - eps_inf)/(1 + j omega tau)
    + sigma/(j omega eps0), to O((omega dt)^2)."""
    sigma, es, ei, tau = mat
    f = 1.0e9
    w = 2 * np.pi * f
    steps = 4000
    series = _e_point(sigma, es, ei, tau, O.DEBYE_FIXED, lambda n: 1e-9 * np.cos(w * n * DT), steps)
    n = np.arange(steps) + 1
    tail = slice(steps // 2, steps)
    A = np.stack([np.cos(w * n[tail] * DT), np.sin(w * n[tail] * DT), np.ones(n[tail].size)], 1)
    cf, *_ = np.linalg.lstsq(A, series[tail], rcond=None)
    e_phasor = cf[0] - 1j * cf[1]
    eps_num = 1e-9 / (EPS0 * e_phasor)
    eps_th = ei + (es - ei) / (1 + 1j * w * tau) + sigma / (1j * w * EPS0)
    assert abs(eps_num - eps_th) / abs(eps_th) < 2e-2


def test_mur_is_textbook_first_order_in_vacuum():
    """With rp = 1 the reference's face formula is Mur's
    E0' = E1 + (S-1)/(S+1) (E1' - E0), S = c dt / h; the x-high Ey face
    uses E1 (level n) in place of E1' (the reference's quirk).  Only face
    cells are written."""
    rng = np.random.default_rng(4)
    shape = (9, 8, 7)
    nx, ny, nz = shape
    g = _grid(shape)
    tgt = [rng.uniform(-1, 1, shape[::-1]).astype(np.float32) for _ in range(3)]
    src = [rng.uniform(-1, 1, shape[::-1]).astype(np.float32) for _ in range(3)]
    before = [t.copy() for t in tgt]
    rps = [np.ones(n, np.float32) for n in (2 * ny * nz, nx * 2 * nz, nx * ny * 2) * 2]
    O.lib().or_update_mur_f32(C.byref(g), *tgt, *src, *rps)
    c = 1.0 / np.sqrt(np.float64(np.float32(MU0 * EPS0)) * 1.0)
    kk = (c * DT / DX - 1) / (c * DT / DX + 1)
    mur = lambda t1, s1, s0: s1 + kk * (t1 - s0)
    ex, ey, ez = [b.astype(np.float64) for b in before]
    sx, sy, sz = [s.astype(np.float64) for s in src]
    expect = [b.astype(np.float64).copy() for b in before]
    I, J, K = slice(0, nx - 1), slice(1, ny - 1), slice(1, nz - 1)
    expect[0][K, 0, I] = mur(ex[K, 1, I], sx[K, 1, I], sx[K, 0, I])
    expect[0][K, -1, I] = mur(ex[K, -2, I], sx[K, -2, I], sx[K, -1, I])
    expect[0][0, J, I] = mur(ex[1, J, I], sx[1, J, I], sx[0, J, I])
    expect[0][-1, J, I] = mur(ex[-2, J, I], sx[-2, J, I], sx[-1, J, I])
    Jy, Iy = slice(0, ny - 1), slice(1, nx - 1)
    expect[1][K, Jy, 0] = mur(ey[K, Jy, 1], sy[K, Jy, 1], sy[K, Jy, 0])
    expect[1][K, Jy, -1] = mur(sy[K, Jy, -2], sy[K, Jy, -2], sy[K, Jy, -1])
    expect[1][0, Jy, Iy] = mur(ey[1, Jy, Iy], sy[1, Jy, Iy], sy[0, Jy, Iy])
    expect[1][-1, Jy, Iy] = mur(ey[-2, Jy, Iy], sy[-2, Jy, Iy], sy[-1, Jy, Iy])
    Kz = slice(0, nz - 1)
    expect[2][Kz, J, 0] = mur(ez[Kz, J, 1], sz[Kz, J, 1], sz[Kz, J, 0])
    expect[2][Kz, J, -1] = mur(ez[Kz, J, -2], sz[Kz, J, -2], sz[Kz, J, -1])
    expect[2][Kz, 0, Iy] = mur(ez[Kz, 1, Iy], sz[Kz, 1, Iy], sz[Kz, 0, Iy])
    expect[2][Kz, -1, Iy] = mur(ez[Kz, -2, Iy], sz[Kz, -2, Iy], sz[Kz, -1, Iy])
    for q in range(3):
        np.testing.assert_allclose(tgt[q], expect[q], rtol=2e-5, atol=2e-6)
        changed = tgt[q] != before[q]
        assert not changed[1:-1, 1:-1, 1:-1].any()  # interior untouched


def test_mur_roles_of_the_rp_planes():
    """With rp varying cell by cell, every face value satisfies Mur's
    first-order one-way wave equation centred between the face cell (0) and
    its neighbour (1) at t^{n+1/2}, each node's time difference taken at its
    own speed v_i = c / sqrt(rp_i):
        dt (E1' + E1 - E0' - E0) = s0 (E0' - E0) + s1 (E1' - E1),
        s_i = h sqrt(mu0 eps0 rp_i).
    That fixes which rp plane is the face's and which the neighbour's
    (fdtd_calculations.cu:237-427); the same values with the roles swapped
    leave residuals thousands of times larger.  The rp planes are filled
    from a per-cell field the way the rp set-up fills them from materials
    (index a is cell a on a low face and cell n-2+a on a high one)."""
    rng = np.random.default_rng(11)
    shape = (9, 8, 7)
    nx, ny, nz = shape
    g = _grid(shape)
    tgt = [rng.uniform(-1, 1, shape[::-1]).astype(np.float32) for _ in range(3)]
    src = [rng.uniform(-1, 1, shape[::-1]).astype(np.float32) for _ in range(3)]
    before = [t.copy() for t in tgt]
    rpc = rng.uniform(1.0, 60.0, shape[::-1]).astype(np.float32)          # (k, j, i)
    rpx0 = np.ascontiguousarray(rpc[:, :, 0:2]).ravel()                   # [k][j][a], cell i = a
    rpxE = np.ascontiguousarray(rpc[:, :, nx - 2:]).ravel()               # cell i = nx - 2 + a
    rpy0 = np.ascontiguousarray(rpc[:, 0:2, :]).ravel()                   # [k][a][i]
    rpyE = np.ascontiguousarray(rpc[:, ny - 2:, :]).ravel()
    rpz0 = np.ascontiguousarray(rpc[0:2]).ravel()                         # [a][j][i]
    rpzE = np.ascontiguousarray(rpc[nz - 2:]).ravel()
    O.lib().or_update_mur_f32(C.byref(g), *tgt, *src, rpx0, rpy0, rpz0, rpxE, rpyE, rpzE)
    m = np.float64(np.float32(MU0)) * np.float64(np.float32(EPS0))
    slow = np.float64(DX) * np.sqrt(m * rpc.astype(np.float64))          # h / v per cell
    dt = np.float64(DT)
    I, J, K = slice(0, nx - 1), slice(1, ny - 1), slice(1, nz - 1)
    Jy, Iy, Kz = slice(0, ny - 1), slice(1, nx - 1), slice(0, nz - 1)
    # (field, face index, neighbour index, E1' from level n+1 or n): every face the update writes
    faces = [(0, (K, 0, I), (K, 1, I), 1), (0, (K, -1, I), (K, -2, I), 1),
             (0, (0, J, I), (1, J, I), 1), (0, (-1, J, I), (-2, J, I), 1),
             (1, (K, Jy, 0), (K, Jy, 1), 1), (1, (K, Jy, -1), (K, Jy, -2), 0),   # x-high Ey: E1 twice
             (1, (0, Jy, Iy), (1, Jy, Iy), 1), (1, (-1, Jy, Iy), (-2, Jy, Iy), 1),
             (2, (Kz, J, 0), (Kz, J, 1), 1), (2, (Kz, J, -1), (Kz, J, -2), 1),
             (2, (Kz, 0, Iy), (Kz, 1, Iy), 1), (2, (Kz, -1, Iy), (Kz, -2, Iy), 1)]
    worst, worst_swapped = 0.0, np.inf
    for q, f0, f1, new in faces:
        e0n, e1n = tgt[q][f0].astype(np.float64), (tgt if new else src)[q][f1].astype(np.float64)
        e0, e1 = src[q][f0].astype(np.float64), src[q][f1].astype(np.float64)
        s0, s1 = slow[f0], slow[f1]
        scale = (dt + s0 + s1) * 2.0
        res = dt * (e1n + e1 - e0n - e0) - s0 * (e0n - e0) - s1 * (e1n - e1)
        worst = max(worst, float(np.max(np.abs(res) / scale)))
        # the same face value with the planes' roles swapped, checked against the same equation
        sw = ((dt - s0) * e1n + (dt + s0) * e1 - (dt - s1) * e0) / (dt + s1)
        res_sw = dt * (e1n + e1 - sw - e0) - s0 * (sw - e0) - s1 * (e1n - e1)
        worst_swapped = min(worst_swapped, float(np.max(np.abs(res_sw) / scale)))
    assert worst < 2e-6, worst
    assert worst_swapped > 1e3 * worst and worst_swapped > 1e-3, (worst, worst_swapped)
    for q in range(3):
        assert not (tgt[q] != before[q])[1:-1, 1:-1, 1:-1].any()  # interior untouched


def _pulse_run(mur, steps, shape=(60, 60, 60)):
    nx, ny, nz = shape
    s = O.Sim(shape, DT, DX, DX, DX, MU0, EPS0, prec=64, mur=mur)
    # derivative of a Gaussian: zero net charge, so no static field stays behind
    t = (np.arange(steps) - 20) / 6.0
    jz = (-2 * t * np.exp(-t * t)).astype(np.float32)
    s.set_sources([[nx // 2, ny // 2, nz // 2]], jz, jz_mode=1)
    energy = []
    for _ in range(steps):
        s.step()
        energy.append(s.energy())
    return np.asarray(energy)


def test_mur_absorbs_outgoing_pulse():
    """A pulse from the centre of a vacuum box leaves through Mur faces;
    with the faces held (no Mur) it stays trapped."""
    e_open = _pulse_run(True, 320)
    e_closed = _pulse_run(False, 320)
    after = 42  # the pulse is over and has not reached a face yet: both boxes hold the same energy
    assert e_open[after] == pytest.approx(e_closed[after], rel=1e-6)
    assert e_open[-1] < 0.05 * e_open[after]           # left through the Mur faces
    assert e_closed[-1] == pytest.approx(e_closed[after], rel=0.05)  # a lossless closed box


def test_sources_linear_and_held_three_steps():
    shape = (16, 15, 14)
    base = np.zeros(12, np.float32)
    base[3] = 0.9
    runs = []
    for amp in (1.0, 2.0):
        s = O.Sim(shape, DT, DX, DX, DX, MU0, EPS0, prec=64)
        s.set_sources([[7, 7, 6]], base * amp, jz_mode=0)
        s.step(6)
        runs.append(s)
    for f in range(9):
        a, b = runs[0].get(f), runs[1].get(f)
        np.testing.assert_allclose(b, 2 * a, rtol=1e-12, atol=1e-30)
    # jz[3*floor(n/3)]: only steps 3, 4, 5 inject jz[3]; the charge left at
    # the source node is -3 jz[3]
    dz = runs[0].get("Dz")
    dx, dy = runs[0].get("Dx"), runs[0].get("Dy")
    x, y, z = 7, 7, 6
    div = (dx[z, y, x] - dx[z, y, x - 1]) + (dy[z, y, x] - dy[z, y - 1, x]) + (dz[z, y, x] - dz[z - 1, y, x])
    assert div == pytest.approx(-3 * float(np.float32(0.9)), rel=1e-9)


def test_centred_source_is_mirror_symmetric():
    """Ez from a centred Jz source is even under x and y reflections about
    the source (a transposed index in any stencil breaks this)."""
    shape = (31, 31, 31)
    s = O.Sim(shape, DT, DX, DX, DX, MU0, EPS0, prec=64)
    jz = np.sin(np.arange(30) * 0.4).astype(np.float32)
    s.set_sources([[15, 15, 15]], jz, jz_mode=1)
    s.step(18)
    ez = s.get("Ez")
    np.testing.assert_allclose(ez, ez[:, :, ::-1], rtol=1e-9, atol=1e-12 * np.abs(ez).max())
    np.testing.assert_allclose(ez, ez[:, ::-1, :], rtol=1e-9, atol=1e-12 * np.abs(ez).max())
    np.testing.assert_allclose(ez, np.transpose(ez, (0, 2, 1)), rtol=1e-9, atol=1e-12 * np.abs(ez).max())
    assert np.abs(ez).max() > 0


def test_fp32_transcription_tracks_fp64_oracle():
    """On the 64^3 Debye-sphere case the FP32 transcription's probe stays
    within 1e-4 rel-L2 of the FP64 oracle (FIXED); the reference's d
    diverges (|E| grows past 1e20 within 200 steps)."""
    shape = (40, 40, 40)
    ids = np.zeros(shape[::-1], np.uint8)
    k, j, i = np.meshgrid(*[np.arange(40)] * 3, indexing="ij")
    ids[(i - 20) ** 2 + (j - 20) ** 2 + (k - 20) ** 2 <= 64] = 1
    table = np.asarray([(0, 1, 1, 0), (0.8, 55, 40, 8e-12)], np.float32)
    t = table[ids.astype(np.int64)]
    jz = (np.exp(-((np.arange(300) - 60) / 20.0) ** 2) * np.cos(np.arange(300) * 0.06)).astype(np.float32)
    probes = [[8, 20, 20, 20], [8, 12, 20, 20]]
    series = {}
    for prec in (32, 64):
        s = O.Sim(shape, DT, DX, DX, DX, MU0, EPS0, prec=prec)
        s.set_materials(t[..., 0], t[..., 1], t[..., 2], t[..., 3])
        s.set_sources([[10, 20, 20]], jz, jz_mode=1)
        series[prec] = s.run(300, probes)
    assert rel_l2(series[32], series[64]) < 1e-4
    s = O.Sim(shape, DT, DX, DX, DX, MU0, EPS0, prec=64, debye=O.DEBYE_REF)
    s.set_materials(t[..., 0], t[..., 1], t[..., 2], t[..., 3])
    s.set_sources([[10, 20, 20]], jz, jz_mode=1)
    s.step(200)
    assert np.abs(s.get("Ez")).max() > 1e20
