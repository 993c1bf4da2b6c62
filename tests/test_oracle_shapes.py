"""Pins for the oracle's blur_points / dilate_points (``_powerfit.pyx:75-206``).

Each check is a closed form over the whole grid (NumPy broadcasting over voxel
coordinates and their periodic images), not the oracle's per-atom box loop."""
import numpy as np
import pytest

from oracle.shapes import blur_points, dilate_points


def _images(n, lo, hi):
    """For each index v in [0, n): the unwrapped coordinates u in {v - n, v}
    with lo <= u <= hi (the reference writes u at index u mod n)."""
    v = np.arange(n)
    return [(u, (u >= lo) & (u <= hi)) for u in (v - n, v)]


def closed_form(points, weights, extent, shape, wraparound, gaussian=True, sigma=1.0):
    out = np.zeros(shape)
    nz, ny, nx = shape
    for k in range(points.shape[1]):
        px, py, pz = points[:, k]
        e = extent[k]
        lim = [(-d + 1) if wraparound else 0 for d in (nx, ny, nz)]
        bx = (max(int(np.ceil(px - e)), lim[0]), min(int(np.floor(px + e)), nx - 1))
        by = (max(int(np.ceil(py - e)), lim[1]), min(int(np.floor(py + e)), ny - 1))
        bz = (max(int(np.ceil(pz - e)), lim[2]), min(int(np.floor(pz + e)), nz - 1))
        for uz, mz in _images(nz, *bz):
            for uy, my in _images(ny, *by):
                for ux, mx in _images(nx, *bx):
                    d2 = ((ux[None, None, :] - px) ** 2 + (uy[None, :, None] - py) ** 2
                          + (uz[:, None, None] - pz) ** 2)
                    sel = mz[:, None, None] & my[None, :, None] & mx[None, None, :] & (d2 <= e * e)
                    if gaussian:
                        out += np.where(sel, weights[k] * np.exp(-d2 / (2 * sigma * sigma)), 0.0)
                    else:
                        out = np.where(sel, 1.0, out)
    return out


def test_single_atom_is_truncated_gaussian():
    shape = (20, 22, 24)
    p = np.array([[9.3], [10.7], [8.1]])
    sigma = 1.3
    out = np.zeros(shape)
    blur_points(p, [2.5], sigma, out, True)
    z, y, x = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    d2 = (x - 9.3) ** 2 + (y - 10.7) ** 2 + (z - 8.1) ** 2
    ref = np.where(d2 <= (4 * sigma) ** 2, 2.5 * np.exp(-d2 / (2 * sigma ** 2)), 0.0)
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=0)


def test_boundary_distance_is_included():
    # 4 sigma = 2 exactly: the voxel at distance 2 on an axis is written
    out = np.zeros((12, 12, 12))
    blur_points(np.array([[5.0], [5.0], [5.0]]), [1.0], 0.5, out, False)
    assert out[5, 5, 7] == pytest.approx(np.exp(-8.0), rel=1e-15)
    assert out[5, 5, 8] == 0.0
    assert np.count_nonzero(out) == np.count_nonzero(
        np.sum((np.indices(out.shape) - 5.0) ** 2, axis=0) <= 4.0)


@pytest.mark.parametrize("wraparound", [False, True])
def test_wrap_below_clip_above(wraparound):
    """Atoms near voxel 0 wrap to the top indices only with wraparound; atoms
    near the top never wrap to the bottom."""
    rng = np.random.default_rng(11)
    shape = (16, 18, 20)
    pts = np.concatenate([rng.uniform(-1.5, 1.5, (3, 6)),
                          np.array([[19.2], [17.4], [15.6]]),
                          rng.uniform(0, 16, (3, 10))], axis=1)
    w = rng.uniform(1, 16, pts.shape[1])
    sigma = 0.9
    out = np.zeros(shape)
    blur_points(pts, w, sigma, out, wraparound)
    ref = closed_form(pts, w, np.full(pts.shape[1], 4 * sigma), shape, wraparound, True, sigma)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-300)
    top = np.zeros(shape)
    blur_points(pts[:, 6:7], w[6:7], sigma, top, wraparound)
    assert top[:, :, :2].max() == 0.0 and top[:, :, -1].max() > 0.0
    low = np.zeros(shape)
    blur_points(pts[:, :1] * 0 + 0.25, w[:1], sigma, low, wraparound)
    assert (low[:, :, -1].max() > 0.0) == wraparound


def test_box_wider_than_grid_adds_both_images():
    # extent 4 sigma = 6 on an 8-voxel axis: index v gets u = v - 8 and u = v
    shape = (8, 8, 8)
    p = np.array([[0.5], [0.5], [0.5]])
    out = np.zeros(shape)
    blur_points(p, [1.0], 1.5, out, True)
    ref = closed_form(p, np.ones(1), np.array([6.0]), shape, True, True, 1.5)
    np.testing.assert_allclose(out, ref, rtol=1e-13)


def test_additive_over_atoms_and_linear_in_weights():
    rng = np.random.default_rng(5)
    shape = (24, 24, 24)
    pts = rng.uniform(-3, 20, (3, 40))
    w = rng.uniform(1, 16, 40)
    a = np.zeros(shape)
    blur_points(pts, w, 1.1, a, True)
    b = np.zeros(shape)
    blur_points(pts[:, :17], w[:17], 1.1, b, True)
    blur_points(pts[:, 17:], w[17:], 1.1, b, True)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)
    c = np.zeros(shape)
    blur_points(pts, 3.0 * w, 1.1, c, True)
    np.testing.assert_allclose(c, 3.0 * a, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("wraparound", [False, True])
def test_dilate_is_union_of_balls(wraparound):
    rng = np.random.default_rng(3)
    shape = (20, 16, 18)
    pts = rng.uniform(-2, 17, (3, 25))
    radii = rng.uniform(0.7, 3.5, 25)
    out = np.zeros(shape)
    dilate_points(pts, radii, out, wraparound)
    ref = closed_form(pts, None, radii, shape, wraparound, gaussian=False)
    assert np.array_equal(out, ref)
    assert set(np.unique(out)) <= {0.0, 1.0}


def test_argument_checks():
    with pytest.raises(ValueError):
        blur_points(np.zeros((2, 3)), np.ones(3), 1.0, np.zeros((4, 4, 4)))
    with pytest.raises(ValueError):
        dilate_points(np.zeros((3, 3)), np.ones(2), np.zeros((4, 4, 4)))
