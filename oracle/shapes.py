"""ORACLE (test infrastructure): ``blur_points`` and ``dilate_points`` restated.

ctypes bindings of ``oracle/shapes.c``, the CPU restatement of
``src/powerfit_em/_powerfit.pyx:75-206`` (Gaussian splat of atoms for the
template, ball dilation for the mask), with the reference's arguments:
``blur_points(points, weights, sigma, out, wraparound=False)`` and
``dilate_points(points, radii, out, wraparound=False)``; ``points`` is (3, n)
in grid units (x, y, z rows), ``out`` a C-contiguous float64 (nz, ny, nx)
grid updated in place.

Parity pins (tests/test_oracle_shapes.py): one atom equals the closed-form
Gaussian on its truncated box; per-voxel brute force over periodic images with
the reference's clipping (lower wrap, no upper wrap); linearity in the weights
and additivity over atoms; the mask equals the union of closed-form balls.
"""
import ctypes

import numpy as np

from .rotate import _load as _load_lib


def _prep(points, out):
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[0] != 3:
        raise ValueError("points must be (3, n)")
    if out.dtype != np.float64 or not out.flags.c_contiguous or out.ndim != 3:
        raise ValueError("out must be a C-contiguous float64 3-D grid")
    return pts


def blur_points(points, weights, sigma, out, wraparound=False):
    lib = _load_lib()
    pts = _prep(points, out)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    if w.shape != (pts.shape[1],):
        raise ValueError("one weight per point")
    dp = ctypes.POINTER(ctypes.c_double)
    lib.oracle_blur_points.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_double, dp,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.oracle_blur_points(pts.ctypes.data_as(dp), w.ctypes.data_as(dp), pts.shape[1],
                           float(sigma), out.ctypes.data_as(dp), *out.shape, int(bool(wraparound)))


def dilate_points(points, radii, out, wraparound=False):
    lib = _load_lib()
    pts = _prep(points, out)
    r = np.ascontiguousarray(radii, dtype=np.float64)
    if r.shape != (pts.shape[1],):
        raise ValueError("one radius per point")
    dp = ctypes.POINTER(ctypes.c_double)
    lib.oracle_dilate_points.argtypes = [dp, dp, ctypes.c_int64, dp,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.oracle_dilate_points(pts.ctypes.data_as(dp), r.ctypes.data_as(dp), pts.shape[1],
                             out.ctypes.data_as(dp), *out.shape, int(bool(wraparound)))
