"""The default launch shape of a context (fdtd_default_launch_shape), on the
CPU: the rule's picks for the grids of the launch-shape sweeps measured on
MI355X (256 compute units), where each pick was the measured best or within
2 % of it (one exception, 200x100x100 fp32, within 6 %), and the extension
fields that override it."""
import pytest

import fdtd_amd as F
from tests.helpers import EPS0, MU0


def _shape(storage, grid, cus=256, **kw):
    return F.default_launch_shape(F.make_params(grid, 3.8e-12, 2e-3, mu0=MU0, eps0=EPS0), storage, cus=cus, **kw)


# (storage, grid) -> (tile_x, kchunk); 4 tile rows throughout.  Measured kernel times (us) at the
# pick / at 4 planes with one cell per lane:
#   fp32 320x320x880 1610 / 1670, 256^3 305 / 313, 192^3 139 / 148, 160^3 95 / 102, 128^3 48 / 54,
#   96^3 and 64^3 best at 4 planes; fp16 256^3 pairs 203 / 267, 320x320x880 1101 / 1251,
#   128^3 pairs 35 / 45, 160^3 77 / 82.
PICKS = [
    (F.STORE_F32, (320, 320, 880), 64, 16), (F.STORE_F32, (256, 256, 256), 64, 8),
    (F.STORE_F32, (192, 192, 192), 64, 8), (F.STORE_F32, (160, 160, 160), 64, 16),
    (F.STORE_F32, (128, 128, 128), 64, 8), (F.STORE_F32, (96, 96, 96), 64, 4),
    (F.STORE_F32, (64, 64, 64), 64, 4), (F.STORE_F32, (297, 522, 1719), 64, 16),
    (F.STORE_F16, (256, 256, 256), 128, 16), (F.STORE_F16, (320, 320, 880), 64, 16),
    (F.STORE_BF16, (128, 128, 128), 128, 4), (F.STORE_F16, (160, 160, 160), 64, 16),
    (F.STORE_F16, (96, 96, 96), 64, 4), (F.STORE_F16, (200, 100, 100), 64, 4),
    (F.STORE_F16, (64, 64, 64), 64, 4),
]


@pytest.mark.parametrize("storage,grid,tile_x,kchunk", PICKS)
def test_default_shape_matches_the_sweep(storage, grid, tile_x, kchunk):
    s = _shape(storage, grid)
    assert (s["tile_x"], s["tile_y"], s["kchunk"]) == (tile_x, 4, kchunk), s
    assert s["lane_cells"] == (2 if tile_x == 128 else 1)


def test_extension_fields_fix_their_parts():
    g = (320, 320, 880)
    assert _shape(F.STORE_F32, g, kchunk=4, tile_y=8) == dict(tile_x=64, tile_y=8, kchunk=4, lane_cells=1)
    assert _shape(F.STORE_F16, g, lane_cells=2)["tile_x"] == 128
    assert _shape(F.STORE_F32, g, tile_x=128)["tile_x"] == 128
    assert _shape(F.STORE_F32, (64, 64, 3))["kchunk"] == 3          # never deeper than the grid


def test_fewer_compute_units_take_deeper_chunks():
    # half the resident blocks: 256^3 fills the GPU at 16 planes
    assert _shape(F.STORE_F32, (256, 256, 256), cus=128)["kchunk"] == 16


def test_bad_input_is_refused():
    with pytest.raises(F.FdtdError, match="code -2"):
        _shape(F.STORE_F32, (2, 64, 64))
    with pytest.raises(F.FdtdError, match="code -1"):
        _shape(F.STORE_F32, (64, 64, 64), tile_y=3)
