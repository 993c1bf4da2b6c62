"""ORACLE (test infrastructure): ``rotate_grid3d`` restated.

Two independent restatements of ``src/powerfit_em/_extensions.c:7-196``:

``rotate_grid3d``      ctypes binding of ``oracle/rotate.c`` (plain C, the
                       reference's loop order and FP64 evaluation order).
``rotate_grid3d_np``   vectorised NumPy version of the same arithmetic, used
                       to cross-check the C one on small grids.

Both take the reference's arguments ``(grid, rotmat, radius, out, nearest)``
and write ``out`` in place, leaving voxels outside the sphere untouched.

Parity pins (tests/test_oracle_rotate.py): the reference's known-answer
cases (``tests/test_extensions.py:10-37``, ``tests/test_powerfitter.py:43-65``);
identity = copy inside the sphere; the 24 cube-group rotations are exact
voxel permutations; trilinear equals ``scipy.ndimage.map_coordinates``
(order 1, grid-wrap) at R^T v; and, where ``oracle/_ref`` has been built from
the reference's own ``_extensions.c``, bit-identical output on random grids
and rotations of the 10-degree set.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "oracle C library not built: run __graft_entry__.build() "
                "(or `make -C oracle`)")
        lib = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_rotate_grid3d.argtypes = [
            dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int,
            dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_rotate_grid3d.restype = None
        _lib = lib
    return _lib


def rotate_grid3d(grid, rotmat, radius, out, nearest):
    """C restatement; same call signature as the reference extension."""
    lib = _load()
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    rotmat = np.ascontiguousarray(rotmat, dtype=np.float64).reshape(9)
    if out.dtype != np.float64 or not out.flags.c_contiguous:
        raise ValueError("out must be C-contiguous float64")
    dp = ctypes.POINTER(ctypes.c_double)
    lib.oracle_rotate_grid3d(
        grid.ctypes.data_as(dp), *grid.shape, rotmat.ctypes.data_as(dp),
        int(radius), out.ctypes.data_as(dp), *out.shape, int(bool(nearest)))


def _c_round(v):
    """C ``round()``: half away from zero, exact for |v| < 2**52."""
    a = np.abs(v)
    r = np.floor(a)
    r = r + ((a - r) >= 0.5)
    return np.copysign(r, v)


def sphere_points(radius):
    """Integer (z, y, x) in the reference's loop order with |v|^2 <= r^2."""
    r = np.arange(-radius, radius + 1)
    z, y, x = np.meshgrid(r, r, r, indexing="ij")
    keep = (z * z + y * y + x * x) <= radius * radius
    return z[keep], y[keep], x[keep]


def rotate_grid3d_np(grid, rotmat, radius, out, nearest):
    """Vectorised NumPy restatement (FP64, reference evaluation order)."""
    R = np.asarray(rotmat, dtype=np.float64).reshape(9)
    gz, gy, gx = grid.shape
    oz, oy, ox = out.shape
    z, y, x = sphere_points(radius)
    zf, yf, xf = z.astype(np.float64), y.astype(np.float64), x.astype(np.float64)
    xs = (R[6] * zf + R[3] * yf) + R[0] * xf
    ys = (R[7] * zf + R[4] * yf) + R[1] * xf
    zs = (R[8] * zf + R[5] * yf) + R[2] * xf
    o = z * (oy * ox) + (z < 0) * (oz * oy * ox) + y * ox + (y < 0) * (oy * ox) \
        + x + (x < 0) * ox
    flat = grid.reshape(-1)
    size = flat.size
    gslice = gy * gx

    def rd(i):
        ok = (i >= 0) & (i < size)
        return np.where(ok, flat[np.clip(i, 0, size - 1)], 0.0)

    if nearest:
        x0 = _c_round(xs).astype(np.int64)
        y0 = _c_round(ys).astype(np.int64)
        z0 = _c_round(zs).astype(np.int64)
        g = z0 * gslice + y0 * gx + x0 + (x0 < 0) * gx + (y0 < 0) * gslice \
            + (z0 < 0) * size
        val = rd(g)
    else:
        x0 = np.floor(xs).astype(np.int64)
        y0 = np.floor(ys).astype(np.int64)
        z0 = np.floor(zs).astype(np.int64)
        x1, y1, z1 = x0 + 1, y0 + 1, z0 + 1
        g = z0 * gslice + y0 * gx + x0 + (x0 < 0) * gx + (y0 < 0) * gslice \
            + (z0 < 0) * size
        dx, dy, dz = xs - x0, ys - y0, zs - z0
        dx1, dy1, dz1 = 1 - dx, 1 - dy, 1 - dz
        wx = -(x1 == 0).astype(np.int64) * gx
        wy = -(y1 == 0).astype(np.int64) * gslice
        wz = -(z1 == 0).astype(np.int64) * size
        c00 = rd(g) * dx1 + rd(g + 1 + wx) * dx
        c10 = rd(g + gx + wy) * dx1 + rd(g + gx + wy + 1 + wx) * dx
        c01 = rd(g + gslice + wz) * dx1 + rd(g + gslice + wz + 1 + wx) * dx
        c11 = rd(g + gslice + gx + wz + wy) * dx1 \
            + rd(g + gslice + gx + wz + wy + 1 + wx) * dx
        c0 = c00 * dy1 + c10 * dy
        c1 = c01 * dy1 + c11 * dy
        val = c0 * dz1 + c1 * dz
    # last write wins, as in the reference's sequential loop
    o_rev = o[::-1]
    _, first_in_rev = np.unique(o_rev, return_index=True)
    keep = len(o) - 1 - first_in_rev
    ok = (o[keep] >= 0) & (o[keep] < out.size)
    out.reshape(-1)[o[keep][ok]] = val[keep][ok]
