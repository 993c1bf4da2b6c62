"""Template and mask generation on the GPU (pf_blur_points / pf_dilate_points)
against the oracle's restatement of ``_powerfit.pyx:75-206``.

Bar: the mask is bit-identical; the template matches to 1e-13 relative (the
only difference is the rounding of exp), with the reference's order of
additions per voxel."""
import numpy as np
import pytest

from oracle.shapes import blur_points as oracle_blur
from oracle.shapes import dilate_points as oracle_dilate
from powerfit_amd import volume
from powerfit_amd.synth import make_chain

pytestmark = pytest.mark.gpu

BLUR_CASES = [  # shape (z, y, x), atoms, sigma (voxels), wraparound
    ((32, 32, 32), 500, 0.9, True),
    ((30, 36, 40), 2000, 1.3, True),
    ((24, 20, 28), 300, 0.7, False),
    ((8, 8, 8), 20, 1.5, True),        # boxes wider than the grid: two images per axis
    ((64, 64, 64), 6000, 1.1, True),
]


def _atoms(shape, n, seed):
    rng = np.random.default_rng(seed)
    hi = np.array(shape[::-1], dtype=np.float64)[:, None]
    pts = rng.uniform(-4.0, 1.0, (3, n)) * (hi / 4.0) + rng.uniform(0, 1, (3, n)) * hi
    return rng, pts


@pytest.mark.parametrize("shape,n,sigma,wrap", BLUR_CASES)
def test_blur_points_matches_oracle(shape, n, sigma, wrap):
    print(f"blur {shape} n={n} sigma={sigma} wrap={wrap}", flush=True)
    rng, pts = _atoms(shape, n, n + len(shape))
    w = rng.uniform(1.0, 16.0, n)
    base = np.where(rng.random(shape) < 0.05, rng.standard_normal(shape), 0.0)  # out += ...
    ref = base.copy()
    oracle_blur(pts, w, sigma, ref, wrap)
    out = base.copy()
    volume.blur_points(pts, w, sigma, out, wrap)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"  max relative error {err:.2e}", flush=True)
    assert err <= 1e-13
    assert np.array_equal(out != base, ref != base)


@pytest.mark.parametrize("shape,n,sigma,wrap", BLUR_CASES)
def test_dilate_points_matches_oracle_exactly(shape, n, sigma, wrap):
    print(f"dilate {shape} n={n} wrap={wrap}", flush=True)
    rng, pts = _atoms(shape, n, 7 * n)
    radii = rng.uniform(0.5, 4.0 * sigma, n)
    base = np.where(rng.random(shape) < 0.05, 0.5, 0.0)   # hit voxels become 1
    ref = base.copy()
    oracle_dilate(pts, radii, ref, wrap)
    out = base.copy()
    volume.dilate_points(pts, radii, out, wrap)
    assert np.array_equal(out, ref)


def test_no_atoms_leaves_grid():
    out = np.arange(4 * 5 * 6, dtype=np.float64).reshape(4, 5, 6)
    volume.blur_points(np.zeros((3, 0)), np.zeros(0), 1.0, out, True)
    volume.dilate_points(np.zeros((3, 0)), np.zeros(0), out, True)
    assert np.array_equal(out, np.arange(4 * 5 * 6, dtype=np.float64).reshape(4, 5, 6))


def test_structure_to_shape_like_cli_template_and_mask():
    """The CLI's template (atomic-number weights, blurred at the map
    resolution) and mask (5 A balls) on a map grid, as powerfit.py:245-255
    builds them."""
    rng = np.random.default_rng(2024)
    xyz, z = make_chain(3000, 20.0, rng)
    target = volume.Volume(np.zeros((48, 50, 52), np.float32), voxelspacing=1.6,
                           origin=(-3.0, 1.5, 2.0))
    xyz = xyz + np.asarray(target.origin).reshape(3, 1) + 0.37   # near voxel 0: wraps
    tmpl = volume.structure_to_shape_like(target, xyz, resolution=6.0, weights=z, shape="vol")
    mask = volume.structure_to_shape_like(target, xyz, resolution=6.0, shape="mask")
    grid = (xyz - np.asarray(target.origin).reshape(3, 1)) / target.voxelspacing
    sigma = (6.0 / (np.sqrt(2.0) * np.pi)) / target.voxelspacing
    ref_t = np.zeros(target.shape)
    oracle_blur(grid, z.astype(np.float64), sigma, ref_t, True)
    ref_m = np.zeros(target.shape)
    oracle_dilate(grid, np.full(xyz.shape[1], 5.0 / target.voxelspacing), ref_m, True)
    assert tmpl.array.dtype == np.float64 and tmpl.shape == target.shape
    assert np.abs(tmpl.array - ref_t).max() <= 1e-13 * np.abs(ref_t).max()
    assert np.array_equal(mask.array, ref_m)
    assert ref_t[-1].max() > 0 and ref_m[:, :, -1].max() == 1.0   # the wrap was exercised
