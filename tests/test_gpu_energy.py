"""FP64 energy accumulation on the device: sum over cells of E.D/2 +
mu0 |H|^2/2 of the newest levels.  With fp32 storage the fields are the
oracle's bit for bit, so the device sum equals the oracle's FP64 sum of the
same fields up to the summation order (bound 1e-12 relative)."""
import numpy as np
import pytest

import fdtd_amd as F
from tests.helpers import EPS0, MU0, oracle_sim, random_materials, random_sources

pytestmark = pytest.mark.gpu

DT = 3.8105118e-12
DX = 2.0e-3


def _jz(n=40):
    t = np.arange(n, dtype=np.float64)
    return (-(t - 12) / 4 * np.exp(-((t - 12) / 4) ** 2)).astype(np.float32)


@pytest.mark.parametrize("shape", [(30, 22, 26), (65, 9, 17)])
def test_energy_series_matches_oracle(shape):
    rng = np.random.default_rng(11)
    ids, table = random_materials(rng, shape, 5)
    src = random_sources(rng, shape, 3)
    jz = _jz()
    ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0))
    ctx.set_materials(ids, table)
    ctx.set_sources(src, jz)
    ctx.set_energy(4, 100)
    ctx.run(40)
    steps, energy = ctx.energy()
    assert list(steps) == list(range(4, 41, 4))
    s = oracle_sim(shape, DT, DX)
    t = table[ids.astype(np.int64)]
    s.set_materials(t[..., 0].copy(), t[..., 1].copy(), t[..., 2].copy(), t[..., 3].copy())
    s.set_sources(src, jz, jz_mode=1)
    want = []
    for _ in range(10):
        s.step(4)
        want.append(s.energy())
    want = np.asarray(want)
    assert (want > 0).all()
    np.testing.assert_allclose(energy, want, rtol=1e-12, atol=0)


def test_energy_decays_after_the_pulse_leaves():
    """Vacuum box with Mur faces: once the pulse has left through the
    absorbing faces the stored energy falls well below its peak."""
    shape = (40, 40, 40)
    ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0))
    ctx.set_sources(np.asarray([[20, 20, 20]], np.int32), _jz())
    ctx.set_energy(5, 100)
    ctx.run(300)
    _, e = ctx.energy()
    assert np.isfinite(e).all()
    assert e[-1] < 0.05 * e.max()


@pytest.mark.parametrize("storage", [F.STORE_F32, F.STORE_F16])
def test_snapshot_restart_is_bit_exact(tmp_path, storage):
    """Run 12 steps, save, run 9 more; a fresh context with the same inputs
    that loads the snapshot and runs 9 steps ends bit-equal."""
    shape = (37, 21, 18)
    rng = np.random.default_rng(5)
    ids, table = random_materials(rng, shape, 4)
    src = random_sources(rng, shape, 2)
    jz = _jz()

    def make():
        ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0), storage=storage)
        ctx.set_materials(ids, table)
        ctx.set_sources(src, jz)
        return ctx

    a = make()
    a.run(12)
    a.save(tmp_path / "state.bin")
    a.run(9)
    b = make()
    b.load(tmp_path / "state.bin")
    assert b.n == 12
    b.run(9)
    for f in range(9):
        for lv in ([0] if 3 <= f < 6 else [0, 1, 2]):
            assert np.array_equal(a.get(f, lv).view(np.uint32), b.get(f, lv).view(np.uint32)), (f, lv)
    other = F.Context(F.make_params((36, 21, 18), DT, DX, mu0=MU0, eps0=EPS0), storage=storage)
    with pytest.raises(F.FdtdError, match="-2"):
        other.load(tmp_path / "state.bin")
    # a corrupt header (H buffer index 7 at byte 44) or a cut-off file is refused, and the
    # context keeps its state
    raw = bytearray((tmp_path / "state.bin").read_bytes())
    bad = bytearray(raw)
    bad[44:48] = (7).to_bytes(4, "little")
    (tmp_path / "bad.bin").write_bytes(bytes(bad))
    (tmp_path / "cut.bin").write_bytes(bytes(raw[:len(raw) // 2]))
    c = make()
    for name in ("bad.bin", "cut.bin"):
        with pytest.raises(F.FdtdError, match="-4"):
            c.load(tmp_path / name)
    assert c.n == 0


@pytest.mark.parametrize("storage", [F.STORE_F16, F.STORE_BF16])
def test_energy_of_16bit_fields(storage):
    """Under 16-bit storage the device folds the scales into its FP64 sum:
    it equals the sum of E.D/2 + mu0 |H|^2/2 over the stored fields read back
    in SI units (each read-back value rounded to fp32, hence 1e-6)."""
    shape = (30, 22, 26)
    rng = np.random.default_rng(12)
    ids, table = random_materials(rng, shape, 5)
    ctx = F.Context(F.make_params(shape, DT, DX, mu0=MU0, eps0=EPS0), storage=storage)
    ctx.set_materials(ids, table)
    ctx.set_sources(random_sources(rng, shape, 3), _jz())
    ctx.set_energy(10, 10)
    ctx.run(30)
    _, energy = ctx.energy()
    f = {q: ctx.get(q).astype(np.float64) for q in range(9)}
    want = 0.5 * sum((f[6 + c] * f[c]).sum() for c in range(3)) + 0.5 * float(np.float32(MU0)) * sum(
        (f[3 + c] ** 2).sum() for c in range(3))
    assert want > 0
    np.testing.assert_allclose(energy[-1], want, rtol=1e-6, atol=0)
