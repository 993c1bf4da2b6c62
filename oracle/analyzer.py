"""ORACLE (test infrastructure): ``Analyzer._watershed`` restated.

``src/powerfit_em/analyzer.py:80-96``: for ``steps`` cutoffs stepping from
the map's maximum down to half of it (computed in the map's dtype), label the
6-connected components of ``corr >= cutoff`` with
``scipy.ndimage.label`` (default cross structure, no wrap) and take
``scipy.ndimage.maximum_position`` of every label; the peaks are the set of
all those positions.  scipy is the reference's own library for this step.

Parity pins (tests/test_oracle_analyzer.py): component count and membership
equal a breadth-first flood fill over face neighbours, and each component's
maximum position equals the brute-force argmax, on small random grids.
"""
import numpy as np
from scipy.ndimage import label, maximum_position


def cutoffs(corr, steps):
    max_cc = corr.max()
    min_cc = 0.5 * max_cc
    stepsize = (max_cc - min_cc) / steps
    cutoff = max_cc
    out = []
    for _ in range(steps):
        cutoff -= stepsize
        out.append(cutoff)
    return out


def component_maxima(corr, cutoff):
    labels, nfeatures = label(corr >= cutoff)
    pos = maximum_position(corr, labels, list(range(1, nfeatures + 1)))
    return [tuple(int(p) for p in q) for q in pos]


def watershed_positions(corr, steps=5):
    positions = []
    for c in cutoffs(corr, steps):
        positions += component_maxima(corr, c)
    return set(positions)
