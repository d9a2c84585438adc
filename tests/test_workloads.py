"""Seeded workloads: reproducible draws, stated ranges, and the vacuum
margin that keeps first-order Mur (rp of the vacuum default) stable."""
import numpy as np

from fdtd_amd import workloads as W


def face_margin(ids):
    k, j, i = np.nonzero(ids)
    nz, ny, nx = ids.shape
    return min(i.min(), nx - 1 - i.max(), j.min(), ny - 1 - j.max(), k.min(), nz - 1 - k.max())


def test_batched_set_ranges_and_margin():
    ws = W.batched_set(count=12)
    again = W.batched_set(count=12)
    for w, v in zip(ws, again):
        assert w.shape == v.shape and np.array_equal(w.ids, v.ids) and np.array_equal(w.jz, v.jz)
        nx, ny, nz = w.shape
        assert 48 <= nx <= 160 and 48 <= ny <= 160 and 48 <= nz <= 224
        assert 500 <= w.steps <= 3000 and 3.0e8 <= w.freq <= 6.0e9
        assert 4 <= w.table.shape[0] <= 12 and w.ids.max() < w.table.shape[0]
        assert (w.table[:, 1] >= w.table[:, 2]).all()          # eps_s >= eps_inf
        if w.ids.any():
            assert face_margin(w.ids) >= W.PHANTOM_MARGIN
        s = w.sources
        assert (s[:, 0] >= 1).all() and (s[:, 0] < nx - 1).all() and (s[:, 2] < nz - 1).all()


def test_served_stream_arrivals_and_mix():
    st = W.served_stream(count=30)
    t = np.array([a for a, _, _ in st])
    assert (np.diff(t) > 0).all()
    # mean inter-arrival of a rate-40 Poisson stream, loosely
    assert 0.005 < t[-1] / len(t) < 0.06
    for _, w, storage in st:
        assert all(32 <= n <= 192 for n in w.shape) and 200 <= w.steps <= 4000
        assert storage in ("f32", "f16")
        if w.ids.any():
            assert face_margin(w.ids) >= W.PHANTOM_MARGIN


def test_box_phantom_margins_and_determinism():
    a = W.box_phantom((64, 48, 40), steps=10)
    b = W.box_phantom((64, 48, 40), steps=10)
    assert np.array_equal(a.ids, b.ids)
    M = W.PHANTOM_MARGIN
    inner = a.ids.copy()
    inner[M:-M, M:-M, M:-M] = 0
    assert not inner.any()                     # nothing within the margin of a face
    assert set(np.unique(a.ids)) <= set(range(5)) and a.ids.max() > 0
    nx, ny, nz = a.shape
    for x, y, z in a.sources:
        assert 1 <= x < nx - 1 and 1 <= y < ny - 1 and 0 <= z < nz - 1
    assert a.table.shape == (5, 4) and len(a.jz) == 10
