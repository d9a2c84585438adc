"""Voxel phantom input: the tissue table and PGM slices into uint8 tissue
ids, in the reference's read order (nx*ny codes in file order, whatever the
image width, fdtd.cu:920-933) and in image order (pixel (ix, iy); air
outside the image).  Checked against the oracle's reader, numpy, and the
real RIKEN slices when the reference checkout is present."""
import os

import numpy as np
import pytest

import fdtd_amd as F
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SPECS = os.path.join(HERE, "golden", "mat_specs_riken")
RIKEN = "/root/reference/data/pgmdata"


def write_pgm(path, img, comments=("# test",)):
    h, w = img.shape
    with open(path, "w") as f:
        f.write("P2\n" + "".join(c + "\n" for c in comments) + f"{w} {h}\n{int(img.max())}\n")
        f.write("\n".join(str(int(v)) for v in img.ravel()) + "\n")


@pytest.mark.parametrize("img_shape,grid", [((9, 13), (6, 5, 4)), ((5, 4), (8, 7, 3)), ((7, 7), (7, 7, 2))])
def test_both_read_orders(tmp_path, img_shape, grid):
    rng = np.random.default_rng(3)
    nx, ny, nz = grid
    h, w = img_shape
    imgs = [rng.integers(0, 95, size=(h, w)) for _ in range(nz)]
    for k, im in enumerate(imgs):
        write_pgm(tmp_path / f"v1_{k + 1:05d}.pgm", im)
    img_ids = F.load_phantom_ids(tmp_path, grid, F.PGM_IMAGE, threads=3)
    for k, im in enumerate(imgs):
        want = np.zeros((ny, nx), np.uint8)
        want[:min(h, ny), :min(w, nx)] = im[:ny, :nx]
        np.testing.assert_array_equal(img_ids[k], want)
    if h * w >= nx * ny:   # the reference order needs nx*ny codes per slice
        ref_ids = F.load_phantom_ids(tmp_path, grid, F.PGM_REFERENCE, threads=2)
        for k in range(nz):
            codes = O.load_pgm_codes(str(tmp_path / f"v1_{k + 1:05d}.pgm"), nx * ny)[0]
            np.testing.assert_array_equal(ref_ids[k].ravel(), codes)
            np.testing.assert_array_equal(ref_ids[k].ravel(), imgs[k].ravel()[:nx * ny])
    else:
        with pytest.raises(F.FdtdError, match="-4"):
            F.load_phantom_ids(tmp_path, grid, F.PGM_REFERENCE)


def test_image_order_header_comments(tmp_path):
    im = np.arange(12).reshape(3, 4) % 95
    write_pgm(tmp_path / "v1_00001.pgm", im, comments=("# made by", "# a scanner"))
    ids = F.load_phantom_ids(tmp_path, (4, 3, 1), F.PGM_IMAGE)
    np.testing.assert_array_equal(ids[0], im)


def test_errors(tmp_path):
    write_pgm(tmp_path / "v1_00001.pgm", np.zeros((4, 4), int))
    with pytest.raises(F.FdtdError, match="-4"):          # slice 2 missing
        F.load_phantom_ids(tmp_path, (4, 4, 2), F.PGM_IMAGE)
    write_pgm(tmp_path / "v1_00002.pgm", np.full((4, 4), 95))
    with pytest.raises(F.FdtdError, match="-2"):          # code 95 > 94
        F.load_phantom_ids(tmp_path, (4, 4, 2), F.PGM_IMAGE)
    with pytest.raises(F.FdtdError, match="-4"):
        F.load_tissue_table(tmp_path / "no_such_table")
    # a code or an image size past the range of any integer type is refused, not wrapped
    (tmp_path / "v1_00002.pgm").write_text("P2\n4 4\n94\n" + "0 " * 15 + "9" * 40 + "\n")
    with pytest.raises(F.FdtdError, match="-2"):
        F.load_phantom_ids(tmp_path, (4, 4, 2), F.PGM_IMAGE)
    with pytest.raises(F.FdtdError, match="-2"):
        F.load_phantom_ids(tmp_path, (4, 4, 2), F.PGM_REFERENCE)
    (tmp_path / "v1_00002.pgm").write_text("P2\n4294967297 1\n94\n0\n")
    with pytest.raises(F.FdtdError, match="-4"):
        F.load_phantom_ids(tmp_path, (4, 4, 2), F.PGM_IMAGE)


def test_tissue_table_matches_oracle():
    table, rows = F.load_tissue_table(SPECS)
    want, _ = O.load_specs(SPECS)
    assert rows == 53
    np.testing.assert_array_equal(table.view(np.uint32), np.asarray(want, np.float32).reshape(95, 4).view(np.uint32))


@pytest.mark.skipif(not os.path.isdir(RIKEN), reason="the reference checkout is not present")
def test_riken_slices():
    """Slice 800 of the RIKEN phantom (265x490) read in image order equals its
    pixel matrix; the reference order at the reference's 100x100 grid reads
    the first 10^4 codes of each slice."""
    path = os.path.join(RIKEN, "v1_00800.pgm")
    tok = open(path).read().split()
    w, h = int(tok[3]), int(tok[4])
    pix = np.asarray(tok[6:6 + w * h], np.int64).reshape(h, w)
    ids = F.load_phantom_ids(RIKEN, (w, h, 800), F.PGM_IMAGE)
    np.testing.assert_array_equal(ids[799], pix)
    ref = F.load_phantom_ids(RIKEN, (100, 100, 800), F.PGM_REFERENCE)
    np.testing.assert_array_equal(ref[799].ravel(), pix.ravel()[:10000])
