"""Pins of the oracle's rotate_grid3d restatement (CPU only)."""
import glob
import importlib.util
import os

import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from oracle.rotate import rotate_grid3d, rotate_grid3d_np, sphere_points
from powerfit_amd.rotations import load_set, quat_to_rotmat, rotations_for_angle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden", "rotate_ref.npz")
IMPLS = [rotate_grid3d, rotate_grid3d_np]


def euler_z(deg):
    a = np.radians(deg)
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])


@pytest.mark.parametrize("rot", IMPLS)
def test_known_answers_extensions(rot):
    """tests/test_extensions.py:10-37 of the reference."""
    grid = np.zeros((4, 5, 6))
    for idx in [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 0, 2), (0, 0, -1), (-1, 0, 0)]:
        grid[idx] = 1
    out = np.zeros_like(grid)
    rot(grid, np.eye(3), 2, out, True)
    assert np.allclose(out, grid)
    out.fill(0)
    rot(grid, np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]]), 2, out, False)
    answer = np.zeros_like(out)
    for idx in [(0, 0, 0), (0, 1, 0), (0, 1, -1), (0, 2, 0), (0, -1, 0), (-1, 0, 0)]:
        answer[idx] = 1
    assert np.allclose(answer, out)


@pytest.mark.parametrize("rot", IMPLS)
def test_known_answers_powerfitter(rot):
    """tests/test_powerfitter.py:43-65: identity and 90 deg about z, rmax = 3."""
    shape = (5, 6, 7)
    template = np.zeros(shape)
    template[0, 0, 0:3] = 1
    out = np.zeros(shape)
    rot(template, np.eye(3), 3, out, False)
    assert np.allclose(out, template)
    out = np.zeros(shape)
    rot(template, euler_z(90), 3, out, False)
    answer = np.zeros(shape)
    answer[0, :3, 0] = 1
    assert np.allclose(out, answer)


def test_identity_copies_sphere():
    rng = np.random.default_rng(1)
    g = rng.standard_normal((10, 11, 12))
    r = 5
    for nearest in (0, 1):
        out = np.zeros_like(g)
        rotate_grid3d(g, np.eye(3), r, out, nearest)
        z, y, x = sphere_points(r)
        sel = (z % 10, y % 11, x % 12)
        assert np.array_equal(out[sel], g[sel])
        inside = np.zeros(g.shape, bool)
        inside[sel] = True
        assert np.all(out[~inside] == 0)


def test_cube_group_is_exact_permutation():
    """The 24 signed permutation matrices move voxels without interpolation."""
    R = quat_to_rotmat(load_set("c48u1.npy")[:, :4])
    rng = np.random.default_rng(2)
    g = rng.standard_normal((9, 9, 9))
    r = 4
    z, y, x = sphere_points(r)
    for Rk in R:
        assert np.allclose(np.abs(Rk).sum(axis=0), 1) and set(np.unique(Rk)) <= {-1, 0, 1}
        out_lin = np.zeros_like(g)
        out_nn = np.zeros_like(g)
        rotate_grid3d(g, Rk, r, out_lin, 0)
        rotate_grid3d(g, Rk, r, out_nn, 1)
        src = np.stack([x, y, z]).T @ Rk   # R^T v in (x, y, z)
        sx, sy, sz = [np.rint(src[:, i]).astype(int) for i in range(3)]
        expect = g[sz % 9, sy % 9, sx % 9]
        got_lin = out_lin[z % 9, y % 9, x % 9]
        assert np.allclose(got_lin, expect, atol=1e-12)
        assert np.array_equal(out_nn[z % 9, y % 9, x % 9], expect)


def test_trilinear_equals_map_coordinates():
    """Inside the sphere, trilinear = scipy map_coordinates(order 1, grid-wrap) at R^T v."""
    R = quat_to_rotmat(load_set("c48n309.npy")[:, :4])
    rng = np.random.default_rng(3)
    g = rng.standard_normal((15, 16, 17))   # every extent > 2r: no write collisions
    r = 7
    z, y, x = sphere_points(r)
    for k in rng.choice(len(R), 12, replace=False):
        out = np.zeros_like(g)
        rotate_grid3d(g, R[k], r, out, 0)
        src = np.stack([x, y, z]).T.astype(float) @ R[k]
        ref = map_coordinates(g, [src[:, 2], src[:, 1], src[:, 0]], order=1, mode="grid-wrap")
        got = out[z % 15, y % 16, x % 17]
        assert np.allclose(got, ref, atol=1e-12)


def test_c_and_numpy_restatements_agree():
    rng = np.random.default_rng(4)
    for setname in ("c48n309.npy", "c600v.npy"):
        R = quat_to_rotmat(load_set(setname)[:, :4])
        for shape in [(8, 8, 8), (7, 9, 11), (12, 10, 9)]:
            g = rng.standard_normal(shape)
            m = rng.integers(0, 3, shape).astype(float)
            for k in rng.choice(len(R), 6, replace=False):
                for nearest, src in ((0, g), (1, m)):
                    a = np.zeros(shape)
                    b = np.zeros(shape)
                    rotate_grid3d(src, R[k], min(shape) // 2, a, nearest)
                    rotate_grid3d_np(src, R[k], min(shape) // 2, b, nearest)
                    assert np.array_equal(a, b)


def test_against_reference_vectors():
    """Golden vectors written by the reference's own compiled _extensions.c.

    Nearest: bit-identical.  Trilinear: the reference is built with
    -ffast-math, which reassociates the interpolation sums, so agreement is
    to a few ulp (the same arithmetic built without -ffast-math matches
    bit for bit)."""
    d = np.load(GOLDEN)
    for c in range(int(d["ncases"])):
        out = np.zeros(d[f"case{c}_out"].shape)
        rotate_grid3d(d[f"case{c}_grid"], d[f"case{c}_rot"], int(d[f"case{c}_radius"]), out,
                      int(d[f"case{c}_nearest"]))
        ref = d[f"case{c}_out"]
        if int(d[f"case{c}_nearest"]):
            assert np.array_equal(out, ref), c
        else:
            assert np.allclose(out, ref, rtol=0, atol=1e-13), c


def _load_ref_ext():
    paths = glob.glob(os.path.join(ROOT, "oracle", "_ref", "_extensions*.so"))
    if not paths:
        return None
    spec = importlib.util.spec_from_file_location("_extensions", paths[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_against_reference_build_random():
    ref = _load_ref_ext()
    if ref is None:
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    R = quat_to_rotmat(load_set("c48n309.npy")[:, :4])
    rng = np.random.default_rng(5)
    for shape in [(16, 16, 16), (17, 18, 19)]:
        g = rng.standard_normal(shape)
        m = rng.integers(0, 4, shape).astype(float)
        for k in rng.choice(len(R), 25, replace=False):
            for nearest, src in ((0, g), (1, m)):
                a = np.zeros(shape)
                b = np.zeros(shape)
                ref.rotate_grid3d(src, R[k], min(shape) // 2, a, nearest)
                rotate_grid3d(src, R[k], min(shape) // 2, b, nearest)
                if nearest:
                    assert np.array_equal(a, b)
                else:
                    assert np.allclose(a, b, rtol=0, atol=1e-13)


@pytest.mark.parametrize("nearest", [False, True])
def test_rotated_support_within_radius_plus_two(nearest):
    """Rotation about voxel 0 preserves |v|: every voxel of the rotated grid
    with |v| > ceil(R) + 2 is zero when the source is zero beyond radius R
    (periodic signed coordinates).  The fused forward stage skips those planes
    and rows, so this bound must hold for trilinear and nearest sampling."""
    rng = np.random.default_rng(77)
    shape = (24, 24, 24)
    idx = [np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n) for n in shape]
    z, y, x = np.meshgrid(*idx, indexing="ij")
    d2 = z * z + y * y + x * x
    allrot = rotations_for_angle(10.0)
    rots = allrot[rng.choice(allrot.shape[0], 24, replace=False)]
    for R in (3.3, 5.0, 7.9):
        grid = np.where(d2 <= R * R, rng.random(shape) + 0.1, 0.0)
        zr = int(np.ceil(np.sqrt(d2[grid != 0].max()))) + 2
        for rot in rots:
            out = np.zeros(shape)
            rotate_grid3d(grid, rot, 12, out, nearest)
            nz = out != 0
            assert nz.any()
            assert d2[nz].max() <= zr * zr
