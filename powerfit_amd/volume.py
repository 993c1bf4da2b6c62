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


# ---------------------------------------------------------------------------
# Map files: CCP4 / MRC (volume.py:228-491), and the grid pipeline the CLI
# runs before the search (powerfit.py:218-233): resample, trim, extend.

# The 1024-byte CCP4/MRC header, word by word (the file format).
_MAP_HEADER = [
    ("n", "i4", 3), ("mode", "i4"), ("nstart", "i4", 3), ("m", "i4", 3),
    ("cell", "f4", 3), ("angles", "f4", 3), ("axes", "i4", 3),
    ("amin", "f4"), ("amax", "f4"), ("amean", "f4"),
    ("ispg", "i4"), ("nsymbt", "i4"), ("lskflg", "i4"),
    ("skwmat", "f4", 9), ("skwtrn", "f4", 3), ("future", "f4", 12),
    ("origin", "f4", 3), ("map", "S4"), ("machst", "u1", 4), ("rms", "f4"),
    ("nlabl", "i4"), ("labels", "S800"),
]
_MODES = {0: "i1", 1: "i2", 2: "f4"}


def _header_dtype(endian):
    return np.dtype([(f[0], endian + f[1], f[2]) if len(f) == 3 else (f[0], endian + f[1])
                     for f in _MAP_HEADER])


def read_map(path, fmt=None):
    """(density (z, y, x), voxelspacing, origin (x, y, z)) of a CCP4/MRC map.

    Endianness from the machine stamp (0x44 little, 0x11 big); modes 0/1/2
    (int8, int16 read as int32, float32 read as float64); rectangular cells
    with equal spacing only (RuntimeError otherwise, as the reference).  The
    origin is nstart * spacing for .ccp4/.map and the header origin for .mrc.
    Any extended header is skipped and permuted axis orders are transposed to
    (z, y, x)."""
    fmt = fmt or str(path).rsplit(".", 1)[-1]
    if fmt not in ("ccp4", "map", "mrc"):
        raise ValueError("Extension of file is not supported.")
    with open(path, "rb") as fh:
        raw = fh.read()
    stamp = raw[212]
    if stamp == 0x44:
        endian = "<"
    elif stamp == 0x11:
        endian = ">"
    else:
        raise RuntimeError("Endiannes is not properly set in file. Check the file format.")
    h = np.frombuffer(raw[:1024], dtype=_header_dtype(endian))[0]
    if np.any(np.abs(h["angles"] - 90.0) > 1e-3):
        raise RuntimeError("Only densities in rectangular boxes are supported.")
    # header columns, rows, sections -> file axes; axes[i] names the x/y/z
    # axis (1, 2, 3) that file axis i runs along
    axes = [int(a) for a in h["axes"]]
    if sorted(axes) != [1, 2, 3]:
        raise RuntimeError("Invalid axis order in map header.")
    counts = [int(c) for c in h["n"]]            # along columns, rows, sections
    grid = [int(c) for c in h["m"]]              # sampling along x, y, z
    spacings = [float(h["cell"][a]) / grid[a] for a in range(3)]
    mean = sum(spacings) / 3.0
    if any(abs(s - mean) > 1e-4 for s in spacings):
        raise RuntimeError("Voxel spacing is not equal in all directions.")
    voxelspacing = spacings[0]
    mode = int(h["mode"])
    if mode not in _MODES:
        raise RuntimeError(f"Map mode {mode} is not supported.")
    start = 1024 + int(h["nsymbt"])
    n = counts[0] * counts[1] * counts[2]
    data = np.frombuffer(raw, dtype=endian + _MODES[mode], count=n, offset=start)
    data = data.reshape(counts[::-1])            # (sections, rows, columns)
    # file axis order (sections, rows, columns) -> (z, y, x)
    perm = [2 - sorted(range(3), key=lambda i: axes[i])[k] for k in (2, 1, 0)]
    data = np.transpose(data, perm)
    if mode == 1:
        data = data.astype(np.int32)
    elif mode == 2:
        data = data.astype(np.float64)
    else:
        data = data.copy()
    if fmt in ("ccp4", "map"):
        nstart = [int(v) for v in h["nstart"]]   # along columns, rows, sections
        xyz_start = [0, 0, 0]
        for i, a in enumerate(axes):
            xyz_start[a - 1] = nstart[i]
        origin = np.asarray([v * voxelspacing for v in xyz_start])
    else:
        origin = [float(v) for v in h["origin"]]
    return data, voxelspacing, origin


def write_map(path, volume, fmt=None):
    """Write ``volume`` as CCP4/MRC (volume.py:399-491): little endian, axes
    1 2 3, mode 0 for int8, 1 for int16/int32, 2 for float32/float64; the
    start indices (ccp4/map) or the origin (mrc) carry the position."""
    fmt = fmt or str(path).rsplit(".", 1)[-1]
    if fmt not in ("ccp4", "mrc", "map"):
        raise ValueError("Format is not recognized. Use ccp4, mrc, or map.")
    arr = np.asarray(volume.array)
    name = arr.dtype.name
    if name == "int8":
        mode = 0
    elif name in ("int16", "int32"):
        mode = 1
    elif name in ("float32", "float64"):
        mode = 2
    else:
        raise TypeError("Data type ({:})is not supported.".format(name))
    nz, ny, nx = arr.shape
    h = np.zeros((), dtype=_header_dtype("<"))
    h["n"] = (nx, ny, nz)
    h["mode"] = mode
    h["nstart"] = ([int(round(v)) for v in volume.start] if fmt in ("ccp4", "map")
                   else (0, 0, 0))
    h["m"] = (nx, ny, nz)
    h["cell"] = volume.dimensions
    h["angles"] = (90.0, 90.0, 90.0)
    h["axes"] = (1, 2, 3)
    h["amin"], h["amax"], h["amean"] = arr.min(), arr.max(), arr.mean()
    h["ispg"] = 1
    h["origin"] = volume.origin if fmt == "mrc" else (0.0, 0.0, 0.0)
    h["map"] = b"MAP "
    h["machst"] = (0x44, 0x41, 0, 0)
    h["rms"] = arr.std()
    h["labels"] = b" " * 800
    with open(path, "wb") as fh:
        fh.write(h.tobytes())
        fh.write(arr.astype("<" + _MODES[mode]).tobytes())


def _fromfile(cls, fid, fmt=None):
    array, voxelspacing, origin = read_map(fid, fmt)
    return cls(array, voxelspacing, origin)


def _tofile(self, fid, fmt=None):
    fmt = fmt or str(fid).rsplit(".", 1)[-1]
    if fmt in ("ccp4", "map", "mrc"):
        write_map(fid, self, fmt)
    else:
        raise RuntimeError("Format is not supported.")


Volume.fromfile = classmethod(_fromfile)
Volume.tofile = _tofile


def resample(volume, factor, order=1):
    """Spline zoom by ``factor`` (scipy.ndimage.zoom), spacing / factor."""
    import warnings
    from scipy.ndimage import zoom
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        array = zoom(volume.array, factor, order=order)
    return Volume(array, volume.voxelspacing / factor, volume.origin)


def trim(volume, cutoff, margin=2):
    """Crop to the slabs holding density above ``cutoff``, plus ``margin``
    voxels on each side; the origin moves with the crop."""
    arr = volume.array
    if arr.max() <= cutoff:
        raise ValueError("Cutoff value should be lower than density max.")
    slices = []
    for axis in range(arr.ndim):
        other = tuple(a for a in range(arr.ndim) if a != axis)
        hot = np.nonzero(arr.max(axis=other) > cutoff)[0]
        lo = max(0, int(hot[0]) - margin)
        hi = min(arr.shape[axis], int(hot[-1]) + 1 + margin)
        slices.append(slice(lo, hi))
    origin = [o + volume.voxelspacing * s.start for o, s in zip(volume.origin, slices[::-1])]
    return Volume(arr[tuple(slices)], volume.voxelspacing, origin)


def extend(volume, shape):
    """Zero-pad to ``shape`` (data kept at the low corner)."""
    out = zeros(shape, volume.voxelspacing, volume.origin)
    out.array[tuple(slice(n) for n in volume.shape)] = volume.array
    return out


def is_multiple2357(num):
    """True when ``num`` has no prime factor other than 2, 3, 5 and 7
    (False for num < 1)."""
    if num < 1:
        return False
    for p in (2, 3, 5, 7):
        while num % p == 0:
            num //= p
    return num == 1


def nearest_multiple2357(num):
    """Smallest n >= num with no prime factor other than 2, 3, 5, 7."""
    n = num
    while not is_multiple2357(n):
        n += 1
    return n
