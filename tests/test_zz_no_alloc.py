"""fdtd_dropin_solve_noalloc (the drop-in solve with its state in the
library's static device array, no device allocation) against
fdtd_dropin_solve, through the C ABI: bit-equal probe series on two calls in
a row, the second on a smaller grid (stale contents of the shared array must
not leak into it), then a grid past the array's size is refused."""
import numpy as np
import pytest

import fdtd_amd as F
from tests.test_gpu_batch import make_case

pytestmark = pytest.mark.gpu


def test_noalloc_solve_equals_dropin_solve():
    cases = [make_case(900, (64, 56, 48), 6, 40, jz_mode=F.JZ_HOLD3, rp=True, nprobe=6, jz_limit=512),
             make_case(901, (23, 19, 17), 4, 25, jz_mode=F.JZ_STEP, rp=False, nprobe=5, jz_limit=512)]
    for inst in cases:   # probes beside the first source, so the series carry signal
        x, y, z = (int(v) for v in inst.src[0])
        inst.probes[:3] = [[8, x, y, z], [2, x, y, z], [6, x, y, z]]
    for q, inst in enumerate(cases):
        print(f"no-alloc solve {q}: {inst.shape}, {inst.steps} steps", flush=True)
        want = F.dropin_solve(inst)
        got = F.dropin_solve_noalloc(inst)
        assert got.shape == want.shape == (inst.steps, inst.nprobe)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), f"call {q} {inst.shape}"
        assert np.isfinite(got).all() and np.abs(got).max() > 0
    # the same small grid again right after: still equal
    again = F.dropin_solve_noalloc(cases[1])
    assert np.array_equal(again.view(np.uint64), F.dropin_solve(cases[1]).view(np.uint64))


def test_noalloc_solve_refuses_what_does_not_fit():
    big = make_case(902, (100, 100, 80), 4, 3, nprobe=1, jz_limit=512)
    with pytest.raises(F.FdtdError, match="-2"):
        F.dropin_solve_noalloc(big)
