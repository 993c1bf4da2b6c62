"""Pins for the oracle's watershed (scipy labelling + maximum positions)."""
from collections import deque

import numpy as np
import pytest

from oracle.analyzer import component_maxima, cutoffs, watershed_positions


def flood_components(mask):
    """Face-connected components by breadth-first search, no wrap."""
    seen = np.zeros(mask.shape, bool)
    comps = []
    for start in zip(*np.nonzero(mask)):
        if seen[start]:
            continue
        seen[start] = True
        q, comp = deque([start]), []
        while q:
            p = q.popleft()
            comp.append(p)
            for ax in range(3):
                for d in (-1, 1):
                    n = list(p)
                    n[ax] += d
                    n = tuple(n)
                    if 0 <= n[ax] < mask.shape[ax] and mask[n] and not seen[n]:
                        seen[n] = True
                        q.append(n)
        comps.append(comp)
    return comps


@pytest.mark.parametrize("seed", range(4))
def test_components_and_maxima_equal_flood_fill(seed):
    rng = np.random.default_rng(seed)
    shape = (9, 11, 13)
    corr = rng.random(shape)
    for c in (0.3, 0.55, 0.8):
        comps = flood_components(corr >= c)
        brute = {max(comp, key=lambda p: corr[p]) for comp in comps}
        got = component_maxima(corr, c)
        assert len(got) == len(comps)
        assert set(got) == brute


def test_cutoffs_follow_map_dtype_and_span():
    corr = np.array([[[0.2, 0.9]]], np.float32)
    cuts = cutoffs(corr, 5)
    assert all(isinstance(c, np.float32) for c in cuts)
    assert cuts[-1] == pytest.approx(0.45, rel=1e-6)
    assert cuts[0] == pytest.approx(0.81, rel=1e-6)


def test_watershed_of_two_blobs():
    z, y, x = np.indices((20, 20, 20))
    corr = (np.exp(-((x - 5) ** 2 + (y - 5) ** 2 + (z - 5) ** 2) / 8.0)
            + 0.8 * np.exp(-((x - 14) ** 2 + (y - 13) ** 2 + (z - 12) ** 2) / 8.0))
    assert watershed_positions(corr) == {(5, 5, 5), (12, 13, 14)}
