"""Helpers of the fitting workflow around the search (``src/powerfit_em/helpers.py``)."""
import errno
import os
from math import sqrt

import numpy as np

from .synth import core_indices as determine_core_indices  # noqa: F401  (helpers.py:26-34)


def mkdir_p(path):
    """``helpers.py:16-23``: create a directory, existing ones accepted."""
    try:
        os.makedirs(path)
    except OSError as exc:
        if not (exc.errno == errno.EEXIST and os.path.isdir(path)):
            raise


def fisher_sigma(mv, fsc):
    """``helpers.py:46-47``: sigma of Fisher's z for a molecular volume mv
    (A^3) at resolution fsc: 1 / sqrt(mv / fsc - 3)."""
    return 1 / sqrt(mv / fsc - 3)


def write_fits_to_pdb(structure, solutions, basename="fit"):
    """``helpers.py:50-64``: one PDB per solution, ``<basename>_<n>.pdb`` from
    n = 1: the structure centred on its mean coordinate, rotated by the
    solution's matrix (columns 6-14, row-major) and moved to its (x, y, z)
    (columns 3-5)."""
    centred = structure.duplicate()
    centred.translate(-centred.coor.mean(axis=1))
    for n, sol in enumerate(solutions, start=1):
        out = centred.duplicate()
        out.rotate(np.asarray([float(v) for v in sol[6:]]).reshape(3, 3))
        out.translate(sol[3:6])
        out.tofile(f"{basename}_{n:d}.pdb")
