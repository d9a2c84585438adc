"""Voxel phantoms: PGM slice directories (the reference's data/pgmdata
layout, v1_%05d.pgm per z plane) to uint8 tissue ids, and a compact packed
form for hosts without the slices.

    python -m fdtd_amd.phantom PGM_DIR OUT.npz [--shape NX NY NZ] [--mode image|reference]

writes the ids (compressed, about 5 MB for the 265x490x1687 RIKEN phantom)
with the grid shape and read order.  load() accepts either a PGM directory
or such a file.
"""
import argparse
import os

import numpy as np

from . import PGM_IMAGE, PGM_REFERENCE, load_phantom_ids

RIKEN_SHAPE = (265, 490, 1687)


def pgm_shape(directory):
    """(width, height, slices) of a PGM slice directory."""
    nz = 0
    while os.path.exists(os.path.join(directory, f"v1_{nz + 1:05d}.pgm")):
        nz += 1
    with open(os.path.join(directory, "v1_00001.pgm")) as f:
        tok = []
        for line in f:
            if line.startswith("#"):
                continue
            tok += line.split()
            if len(tok) >= 3:
                break
    return int(tok[1]), int(tok[2]), nz


def pack(directory, out, shape=None, mode=PGM_IMAGE):
    shape = tuple(shape) if shape else pgm_shape(directory)
    ids = load_phantom_ids(directory, shape, mode)
    np.savez_compressed(out, ids=ids, shape=np.asarray(shape, np.int64), mode=np.int64(mode))
    return ids


def load(path, shape=None, mode=PGM_IMAGE):
    """uint8 ids (nz, ny, nx) from a PGM directory or a packed .npz."""
    if os.path.isdir(path):
        return load_phantom_ids(path, tuple(shape) if shape else pgm_shape(path), mode)
    with np.load(path, allow_pickle=False) as z:
        return z["ids"].copy()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("pgm_dir")
    ap.add_argument("out")
    ap.add_argument("--shape", type=int, nargs=3)
    ap.add_argument("--mode", choices=["image", "reference"], default="image")
    a = ap.parse_args()
    ids = pack(a.pgm_dir, a.out, a.shape, PGM_IMAGE if a.mode == "image" else PGM_REFERENCE)
    print(f"{a.out}: ids {ids.shape[::-1]} (nx, ny, nz), {len(np.unique(ids))} tissue codes")


if __name__ == "__main__":
    main()
