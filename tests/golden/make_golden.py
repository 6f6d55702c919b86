"""Writes tests/golden/*.npz: inputs and outputs of the reference's own CPU
kernels (census.cpp, sgm.cpp, postprocess.cpp, compiled unmodified into
oracle/_ref by `make -C oracle ref`) on tiny seeded rigs.

Run in a checkout that has the reference sources:  python tests/golden/make_golden.py
Each fixture holds the five views, D, every intermediate of the helper-class
pipeline for the smallest rig, and for all rigs the API maps, the integer
disparity, the masks and SHA-256 digests of the larger volumes.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sister_amd import synth  # noqa: E402

RIGS = [(64, 48, 16, 101), (96, 64, 24, 102), (128, 96, 32, 103), (80, 60, 72, 104)]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    O.build(ref=True)
    for W, H, D, seed in RIGS:
        rig = synth.rig_small(W, H, D, seed)
        r = O.compute(rig["views"], D, modes=7, debug=True, use_ref=True)
        data = {"views": np.stack(rig["views"]), "D": np.int32(D)}
        for k in ("multiview", "horizontal", "vertical", "disp"):
            data[k] = r[k]
        data["mask"] = np.stack(r["mask"])
        full = (W, H, D) == RIGS[0][:3]
        for F in range(4):
            for img in range(2):
                a = r["census"][2 * F + img]
                if full:
                    data[f"census_{F}_{img}"] = a
                data[f"sha_census_{F}_{img}"] = np.array(digest(a))
            for nm in ("wta_l", "wta_r", "med_l", "med_r"):
                a = r[nm][F].astype(np.uint16)
                data[f"{nm}_{F}"] = a
        for m in range(3):
            for nm in ("msum", "sum"):
                a = r[nm][m]
                if full and m == 0:
                    data[f"{nm}_{m}"] = a
                data[f"sha_{nm}_{m}"] = np.array(digest(a))
        path = os.path.join(HERE, f"ref_{W}x{H}_d{D}.npz")
        np.savez_compressed(path, **data)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
