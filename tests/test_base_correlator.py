"""BaseCorrelator state machine (reference tests/test_powerfitter.py:86-169)."""
import numpy as np
import pytest
from scipy.ndimage import laplace

from powerfit_amd.correlator import BaseCorrelator
from oracle import correlator as oc


@pytest.fixture
def corr():
    return BaseCorrelator(np.random.default_rng(0).random((8, 9, 10)))


def test_init(corr):
    assert corr._target.max() == 1
    assert corr._rmax == 4


def test_rotations_setter(corr):
    corr.rotations = [0] * 9
    assert isinstance(corr._rotations, np.ndarray)
    with pytest.raises(ValueError):
        corr.rotations = [0] * 3
    corr.rotations = [0] * 27
    assert corr._rotations.shape == (3, 3, 3)


def test_template_setter(corr):
    with pytest.raises(ValueError):
        corr.template = np.random.rand(3, 3, 3)
    corr._mask = 1
    corr.template = np.random.rand(8, 9, 10)
    assert corr._mask is None


def test_mask_setter(corr):
    with pytest.raises(ValueError):
        corr.mask = None
    corr._template = np.random.rand(8, 9, 10)
    with pytest.raises(ValueError):
        corr.mask = np.random.rand(10)
    with pytest.raises(ValueError):
        corr.mask = np.zeros((8, 9, 10))
    corr._laplace = True
    t = np.random.rand(8, 9, 10)
    corr._template = t.copy()
    corr.mask = np.ones((8, 9, 10))
    assert np.allclose(corr._mask, 1)
    # template equals the oracle's preparation (Laplace wrap, normalised)
    P, W, N = oc.prepare_template(t, np.ones((8, 9, 10)), laplace=True)
    assert np.allclose(corr._template, P, atol=1e-12)
    assert corr._norm_factor == N


def test_laplace_filter_is_wrap(corr):
    a = np.random.rand(8, 9, 10)
    assert np.allclose(corr._laplace_filter(a), laplace(a, mode="wrap"))


def test_scan_requires_setup(corr):
    with pytest.raises(ValueError):
        corr.scan()
