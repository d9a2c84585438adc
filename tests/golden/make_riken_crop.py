"""Writes tests/golden/riken_crop/: six slices of the reference's RIKEN voxel
phantom (data/pgmdata/v1_00548.pgm .. v1_00553.pgm, 265x490, 53 tissue
codes), each cropped to the 100x100 window at rows 200-299, columns 125-224
(12 tissue codes in slice 550, the most of any 100x100 window of the
trunk slices), renumbered v1_00001.pgm .. v1_00006.pgm and written in the
reference's slice format (P2, the "# test" comment line, width and height,
the largest code, one code per line: fdtd.cu:920-926 skips exactly six
header tokens).  The crop is data, not code: tissue codes only.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_riken_crop.py
"""
import os

import numpy as np

SRC = "/root/reference/data/pgmdata"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "riken_crop")
FIRST, COUNT = 548, 6
ROW0, COL0, SIZE = 200, 125, 100


def read_slice(path):
    tok = open(path).read().split()
    w, h = int(tok[3]), int(tok[4])
    return np.asarray(tok[6:6 + w * h], np.int64).reshape(h, w)


def main():
    os.makedirs(OUT, exist_ok=True)
    for q in range(COUNT):
        img = read_slice(os.path.join(SRC, f"v1_{FIRST + q:05d}.pgm"))[ROW0:ROW0 + SIZE, COL0:COL0 + SIZE]
        with open(os.path.join(OUT, f"v1_{q + 1:05d}.pgm"), "w") as f:
            f.write(f"P2\n# test\n{SIZE} {SIZE}\n{int(img.max())}\n")
            f.write("\n".join(str(int(v)) for v in img.ravel()) + "\n")


if __name__ == "__main__":
    main()
