"""Constructed inputs for the take-best edge cases (shared by CPU and GPU tests)."""
import numpy as np

from powerfit_amd.rotations import rotations_for_angle


def cube_group(angle=10.0):
    """The 24 signed-permutation rotations present in a rotation set, with
    their indices in it (entries exactly 0 or +-1)."""
    R = rotations_for_angle(angle)
    ok = np.all(np.isin(R, (-1.0, 0.0, 1.0)), axis=(1, 2))
    return R[ok], np.flatnonzero(ok)


def zero_variance_case(n, radius=5):
    """Constant target, 0/1 ball mask about voxel 0 with an even voxel count,
    and a template of alternating +1/-1 over the mask.

    With a constant target F[T] and F[T^2] are V at k = 0 and exactly zero
    elsewhere, so for every rotation that permutes voxels (the cube group)
    the FFTs are exact: the masked sum of the template is exactly 0 and so
    is G everywhere, while N<W^2 T^2> - <W T>^2 = N * N - N^2 = 0 exactly
    (V a power of two).  calc_lcc then divides 0 by 0 at every voxel.
    """
    z = np.arange(n)
    s = np.where(z <= n // 2, z, z - n)
    d2 = s[:, None, None] ** 2 + s[None, :, None] ** 2 + s[None, None, :] ** 2
    mask = (d2 <= radius * radius).astype(np.float64)
    idx = np.flatnonzero(mask)
    if idx.size % 2:
        mask.flat[idx[-1]] = 0.0
        idx = idx[:-1]
    template = np.zeros_like(mask)
    template.flat[idx] = np.where(np.arange(idx.size) % 2 == 0, 1.0, -1.0)
    target = np.ones((n, n, n))
    return target, template, mask
