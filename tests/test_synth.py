"""Synthetic workload generator: determinism and basic properties."""
import numpy as np

from powerfit_amd.synth import Workload, make_case, splat


def test_deterministic():
    w = Workload("t", 24, 2.5, 120, 7.0, 20.0, seed=11)
    a, b = make_case(w), make_case(w)
    assert np.array_equal(a.target, b.target) and np.array_equal(a.template, b.template)
    assert a.target.dtype == np.float32 and a.template.dtype == np.float64


def test_splat_total_weight_and_wrap():
    shape = (20, 20, 20)
    p = np.array([[0.3], [19.6], [10.2]])          # near the x/y faces: wraps
    g = splat(p, np.array([6.0]), 1.0, shape)
    # discrete Gaussian sum ~ w (2 pi sigma^2)^(3/2) for sigma = 1 voxel
    assert abs(g.sum() / (6.0 * (2 * np.pi) ** 1.5) - 1) < 0.01
    assert g[10, 19, 0] > 0 and g[10, 0, 19] > 0


def test_mask_and_core_weights():
    w = Workload("t", 24, 2.5, 120, 7.0, 20.0, core_weighted=True, seed=12)
    c = make_case(w)
    assert c.mask.max() >= 2 and np.all(c.mask == np.round(c.mask))
