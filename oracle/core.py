"""ORACLE (test infrastructure): core-weighted mask in closed form.

``determine_core_indices`` (``src/powerfit_em/helpers.py:26-34``) counts how
many successive binary erosions (scipy's default 6-connected cross, zero
border) each voxel survives, plus one.  A voxel survives k erosions exactly
when every voxel within taxicab distance k is inside the mask, with the
outside of the array counting as background; so the count equals the taxicab
distance to the nearest background voxel.  This module computes that
distance directly (exact chamfer transform of the zero-padded mask) and by
brute force for tiny grids.
"""
import itertools

import numpy as np
from scipy.ndimage import distance_transform_cdt


def core_indices(mask):
    fg = np.pad(np.asarray(mask) > 0, 1)
    d = distance_transform_cdt(fg, metric="taxicab")
    return d[1:-1, 1:-1, 1:-1].astype(np.float64)


def core_indices_bruteforce(mask):
    fg = np.asarray(mask) > 0
    padded = np.pad(fg, 1)
    bg = np.argwhere(~padded)
    out = np.zeros(fg.shape)
    for idx in itertools.product(*[range(n) for n in fg.shape]):
        if fg[idx]:
            p = np.array(idx) + 1
            out[idx] = np.abs(bg - p).sum(axis=1).min()
    return out
