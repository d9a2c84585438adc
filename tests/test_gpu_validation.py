"""Range checks of a live context: out-of-range sources, probes, material
ids, field indices, slice planes and halo directions are rejected with
FDTD_ERANGE (-2) or FDTD_EINVAL (-1), and a rejected call leaves the context
as it was (the next steps equal those of a context that never saw it)."""
import numpy as np
import pytest

import fdtd_amd as F
from tests.helpers import EPS0, MU0, random_materials

pytestmark = pytest.mark.gpu

SHAPE = (20, 16, 12)


def _ctx(ids, table, src, jz):
    c = F.Context(F.make_params(SHAPE, 3.8e-12, 2e-3, mu0=MU0, eps0=EPS0))
    c.set_materials(ids, table)
    c.set_sources(src, jz)
    return c


def test_rejected_calls_leave_the_context_unchanged():
    print("validation on a live context", flush=True)
    nx, ny, nz = SHAPE
    rng = np.random.default_rng(31)
    ids, table = random_materials(rng, SHAPE)
    src = np.asarray([[10, 8, 6]], np.int32)
    jz = rng.uniform(-1, 1, 20).astype(np.float32)
    a = _ctx(ids, table, src, jz)
    a.set_probes(np.asarray([[8, 10, 8, 6]], np.int32), 8)
    bad = [
        (lambda: a.set_sources(np.asarray([[0, 8, 6]], np.int32), jz), -2),        # x = 0: updateSources would read Hy(x-1) outside the row
        (lambda: a.set_sources(np.asarray([[10, 8, nz]], np.int32), jz), -2),       # z = nz
        (lambda: a.set_probes(np.asarray([[9, 1, 1, 1]], np.int32), 8), -2),        # field 9
        (lambda: a.set_probes(np.asarray([[8, nx, 1, 1]], np.int32), 8), -2),       # i = nx
        (lambda: a.set_materials(np.full(ids.shape, len(table), np.uint8), table), -2),  # id >= nmat
        (lambda: a.set_materials(ids, np.zeros((129, 4), np.float32)), -2),          # more than 128 materials
        (lambda: a.get(9), -1),
        (lambda: a.get(6, 3), -1),
        (lambda: a.set_slices(nz, 5, 4), -2),
        (lambda: a.halo_bytes(2), -1),
    ]
    for call, code in bad:
        with pytest.raises(F.FdtdError, match=rf"code {code}:"):
            call()
    b = _ctx(ids, table, src, jz)
    b.set_probes(np.asarray([[8, 10, 8, 6]], np.int32), 8)
    a.run(8)
    b.run(8)
    for f in range(9):
        assert np.array_equal(a.get(f).view(np.uint32), b.get(f).view(np.uint32)), F.FIELDS[f]
    assert np.array_equal(a.probes().view(np.uint64), b.probes().view(np.uint64))
    assert np.abs(a.get(8)).max() > 0
