"""Maps and the template/mask generators, with the reference's interface.

Mirrors ``src/powerfit_em/volume.py``: ``Volume`` (:13-56), ``zeros`` /
``zeros_like`` (:58-63), ``res_to_sigma`` / ``sigma_to_res`` (:121-126),
``structure_to_shape`` (:144-189) and ``structure_to_shape_like``
(:192-224), and ``_powerfit.blur_points`` / ``dilate_points``
(``_powerfit.pyx:75-206``), which here run on the GPU through
``pf_blur_points`` / ``pf_dilate_points``.  Map file I/O is not part of this
package.
"""
import numpy as np

from . import _lib


class Volume:
    """Density array with voxel spacing and origin (x, y, z, in Angstrom)."""

    def __init__(self, array, voxelspacing=1.0, origin=(0, 0, 0)):
        self.array = array
        self.voxelspacing = voxelspacing
        self.origin = origin

    @property
    def shape(self):
        return self.array.shape

    @property
    def dimensions(self):
        return np.asarray([x * self.voxelspacing for x in self.array.shape][::-1])

    @property
    def start(self):
        return np.asarray([x / self.voxelspacing for x in self.origin])

    def duplicate(self):
        return Volume(self.array.copy(), voxelspacing=self.voxelspacing, origin=self.origin)


def zeros(shape, voxelspacing, origin):
    return Volume(np.zeros(shape), voxelspacing, origin)


def zeros_like(volume):
    return Volume(np.zeros_like(volume.array), volume.voxelspacing, volume.origin)


def res_to_sigma(resolution):
    return resolution / (np.sqrt(2.0) * np.pi)


def sigma_to_res(sigma):
    return sigma * (np.sqrt(2.0) * np.pi)


def _grid_args(points, out):
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[0] != 3:
        raise ValueError("points must be a (3, n) array")
    if not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.ndim != 3 \
            or not out.flags.c_contiguous or not out.flags.writeable:
        raise ValueError("out must be a writeable C-contiguous float64 3-D array")
    return pts


def blur_points(points, weights, sigma, out, wraparound=False, device=0):
    """Add every atom's Gaussian, weights[k] exp(-d^2 / (2 sigma^2)) within
    4 sigma, to ``out`` (grid units; ``_powerfit.pyx:75-140``)."""
    pts = _grid_args(points, out)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    if w.shape != (pts.shape[1],):
        raise ValueError("weights array is of incorrect size")
    lib = _lib.load()
    _lib.check(lib.pf_blur_points(int(device), _lib.ptr(pts, _lib.DP), _lib.ptr(w, _lib.DP),
                                  pts.shape[1], float(sigma), _lib.ptr(out, _lib.DP),
                                  *out.shape, int(bool(wraparound))), "pf_blur_points")


def dilate_points(points, radii, out, wraparound=False, device=0):
    """Set ``out`` to 1 within radii[k] of every atom (grid units;
    ``_powerfit.pyx:143-206``)."""
    pts = _grid_args(points, out)
    r = np.ascontiguousarray(radii, dtype=np.float64)
    if r.shape != (pts.shape[1],):
        raise ValueError("radii array is of incorrect size")
    lib = _lib.load()
    _lib.check(lib.pf_dilate_points(int(device), _lib.ptr(pts, _lib.DP), _lib.ptr(r, _lib.DP),
                                    pts.shape[1], _lib.ptr(out, _lib.DP), *out.shape,
                                    int(bool(wraparound))), "pf_dilate_points")


def _shape_inputs(n, voxelspacing, radii, weights, shape):
    if shape == "vol":
        if weights is None:
            weights = np.ones(n)
        elif np.asarray(weights).size != n:
            raise ValueError("weights array is of incorrect size")
    if shape == "mask":
        if radii is None:
            radii = np.empty(n, dtype=np.float64)
            radii.fill(5)
        elif np.asarray(radii).size != n:
            raise ValueError("weights array is of incorrect size")
        radii = np.asarray(radii, dtype=np.float64) / voxelspacing
    return radii, weights


def structure_to_shape(xyz, resolution, out=None, voxelspacing=None, radii=None, weights=None,
                       shape="vol", device=0):
    """Template ('vol') or mask ('mask') of atoms ``xyz`` (3, n) in Angstrom,
    on ``out``'s grid or on a new grid that holds the atoms
    (``volume.py:144-189``)."""
    if shape not in ("vol", "mask"):
        raise ValueError("shape should either be 'vol' or 'mask'")
    xyz = np.asarray(xyz, dtype=np.float64)
    if out is None and voxelspacing is None:
        voxelspacing = resolution / 4.0
    else:
        voxelspacing = out.voxelspacing
    radii, weights = _shape_inputs(xyz.shape[1], voxelspacing, radii, weights, shape)
    sigma = res_to_sigma(resolution)
    if out is None:
        extend = 4 * sigma
        imin = xyz.min(axis=1) - extend
        imax = xyz.max(axis=1) + extend
        center = (imin + imax) / 2.0
        ni = (np.ceil((imax - imin) / voxelspacing)).astype(np.int32)
        origin = center - (ni * voxelspacing) / 2.0
        grid = np.zeros(ni[::-1])
        xyz_grid = xyz + (ni * voxelspacing / 2.0 - center).reshape(-1, 1)
        out = Volume(grid, voxelspacing, origin)
    else:
        xyz_grid = xyz - np.asarray(out.origin, dtype=np.float64).reshape(-1, 1)
    xyz_grid = xyz_grid / voxelspacing
    if shape == "vol":
        blur_points(xyz_grid, weights, sigma / voxelspacing, out.array, True, device)
    else:
        dilate_points(xyz_grid, radii, out.array, True, device)
    return out


def structure_to_shape_like(vol, xyz, resolution=None, weights=None, radii=None, shape="vol",
                            device=0):
    """Template ('vol', atoms blurred at ``resolution`` with ``weights``) or
    mask ('mask', balls of ``radii`` Angstrom, 5 by default) of atoms ``xyz``
    (3, n) in Angstrom on the grid of ``vol`` (``volume.py:192-224``: the
    CLI's template and mask, ``powerfit.py:245-255``).  Unit weights when none
    are given."""
    if resolution is None:
        resolution = vol.resolution
    xyz = np.asarray(xyz, dtype=np.float64)
    radii, weights = _shape_inputs(xyz.shape[1], vol.voxelspacing, radii, weights, shape)
    sigma = (resolution / (np.sqrt(2.0) * np.pi)) / vol.voxelspacing
    xyz_grid = xyz - np.asarray(vol.origin, dtype=np.float64).reshape(3, 1)
    xyz_grid /= vol.voxelspacing
    out = Volume(np.zeros(vol.shape), vol.voxelspacing, vol.origin)   # FP64 grid
    if shape == "vol":
        blur_points(xyz_grid, weights, sigma, out.array, True, device)
    elif shape == "mask":
        dilate_points(xyz_grid, radii, out.array, True, device)
    return out
