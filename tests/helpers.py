"""Shared helpers of the test suite: seeded random states and bit-level
comparison.  The oracle is imported here only as the checker."""
import numpy as np

from oracle import oracle as O

EPS0 = 8.841941e-12
MU0 = 1.2566371e-06


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def first_diff(a, b):
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    if bad.size == 0:
        return "equal"
    q = bad[0]
    return f"{bad.size} cells differ; first flat index {q}: {a[q]!r} vs {b[q]!r}"


def rel_l2(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / den) if den > 0 else float(np.linalg.norm(x - ref))


def random_materials(rng, shape, nmat=5):
    nx, ny, nz = shape
    table = np.zeros((nmat, 4), np.float32)
    table[0] = (0.0, 1.0, 1.0, 0.0)
    for m in range(1, nmat):
        ei = rng.uniform(4, 45)
        es = rng.uniform(ei, 80)
        table[m] = (rng.uniform(0.02, 2.5), es, ei, rng.uniform(5e-12, 50e-12))
    ids = rng.integers(0, nmat, size=(nz, ny, nx)).astype(np.uint8)
    return ids, table


def random_rp(rng, shape):
    nx, ny, nz = shape
    return [rng.uniform(1.0, 60.0, size=n).astype(np.float32) for n in (2 * ny * nz, nx * 2 * nz, nx * ny * 2) * 2]


def random_sources(rng, shape, count):
    nx, ny, nz = shape
    return np.stack([rng.integers(1, nx - 1, count), rng.integers(1, ny - 1, count),
                     rng.integers(0, nz - 1, count)], 1).astype(np.int32)


def random_state(rng, shape):
    """{(field, level): array} for E and D levels 0..2 and H (level 0)."""
    nx, ny, nz = shape
    st = {}
    for f in range(9):
        levels = [0] if 3 <= f < 6 else [0, 1, 2]
        scale = 1.0 if f >= 6 else (EPS0 * 10 if f < 3 else 1.0 / 377.0)
        for lv in levels:
            st[(f, lv)] = (rng.uniform(-1, 1, size=(nz, ny, nx)) * scale).astype(np.float32)
    return st


def oracle_sim(shape, dt, dx, prec=32, debye=O.DEBYE_FIXED):
    return O.Sim(shape, dt, dx, dx, dx, MU0, EPS0, prec=prec, debye=debye)
