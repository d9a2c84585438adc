"""Kernel timing of a context (fdtd_ctx_set_timing / fdtd_ctx_kernel_time):
every `every`-th fused launch is timed by events bound to the dispatch, at
most `capacity` of them, and timed launches compute what untimed ones do."""
import numpy as np
import pytest

import fdtd_amd as F
from tests.helpers import EPS0, MU0, random_materials

pytestmark = pytest.mark.gpu

SHAPE = (70, 40, 30)


def _ctx(seed):
    rng = np.random.default_rng(seed)
    ids, table = random_materials(rng, SHAPE)
    c = F.Context(F.make_params(SHAPE, 3.8e-12, 2e-3, mu0=MU0, eps0=EPS0))
    c.set_materials(ids, table)
    c.set_sources(np.asarray([[20, 20, 15]], np.int32), np.sin(np.arange(40) * 0.3).astype(np.float32))
    return c


def test_timing_counts_and_leaves_results_alone():
    print("timing on", SHAPE, flush=True)
    a, b = _ctx(7), _ctx(7)
    a.set_timing(4, every=3)
    a.run(20)   # launches 0, 3, 6, 9 are timed; capacity 4 stops there
    b.run(20)
    ms, n = a.kernel_time()
    assert n == 4
    assert 0.0 < ms / n < 50.0   # a 84k-cell step takes microseconds
    for f in range(9):
        assert np.array_equal(a.get(f).view(np.uint32), b.get(f).view(np.uint32)), F.FIELDS[f]
    # timing again starts a new count
    a.set_timing(10, every=1)
    a.run(5)
    ms, n = a.kernel_time()
    assert n == 5 and ms > 0.0
