"""Argument checks that run before any device work: bad extension values and
grid sizes are rejected with FDTD_EINVAL (-1) or FDTD_ERANGE (-2) and a
message naming the value, with or without a GPU present."""
import pytest

import fdtd_amd as F
from tests.helpers import EPS0, MU0


def _params(shape):
    return F.make_params(shape, 1e-12, 1e-3, mu0=MU0, eps0=EPS0)


@pytest.mark.parametrize("kw,word", [
    (dict(tile_y=3), "tile_y"),
    (dict(tile_x=96), "tile_x"),
    (dict(lane_cells=3), "lane_cells"),
    (dict(storage=9), "storage"),
    (dict(xcd_remap=5), "xcd_remap"),
    (dict(xcd_remap=-1), "xcd_remap"),
    (dict(time_steps=3), "time_steps"),
    (dict(time_steps=-1), "time_steps"),
    (dict(kchunk=-1), "kchunk"),
    (dict(jz_limit=-2), "jz_limit"),
    (dict(debye=7), "debye_mode"),
    (dict(jz_mode=4), "jz_index_mode"),
    (dict(scales=(-1.0, 0.0, 0.0)), "scale"),
])
def test_bad_extension_rejected(kw, word):
    with pytest.raises(F.FdtdError, match=rf"code -1: .*{word}"):
        F.Context(_params((8, 8, 8)), **kw)


@pytest.mark.parametrize("shape", [(2, 8, 8), (8, 2, 8), (8, 8, 2), (0, 5, 5)])
def test_grid_edges_below_three_rejected(shape):
    with pytest.raises(F.FdtdError, match=r"code -2: grid"):
        F.Context(_params(shape))


def test_grid_beyond_32bit_offsets_rejected():
    """The fused kernel addresses each field array (with its guards) with 32-bit
    byte offsets; a grid of 2^30 elements or more is refused up front."""
    with pytest.raises(F.FdtdError, match=r"code -2: .*too large"):
        F.Context(_params((2048, 2048, 300)))


def test_extension_defaults():
    """The library's defaults as the Python layer sees them: the row block
    order (xcd_remap 3) and one step per launch (time_steps 0)."""
    e = F.make_ext()
    assert e.xcd_remap == 3 and e.time_steps == 0 and e.load_zero_terms == 0 and e.storage == F.STORE_F32
    assert F.make_ext(xcd_remap=0, time_steps=2).xcd_remap == 0
