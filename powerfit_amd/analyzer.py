"""``Analyzer``: peak extraction from the search result, reference interface.

Mirrors ``src/powerfit_em/analyzer.py``: ``Analyzer(corr, rotmat, rotmat_ind,
steps=5, voxelspacing=1, origin=(0, 0, 0), z_sigma=1)``; setting ``steps``
runs the watershed (:80-96); ``solutions`` (:56-78) lists, best first,
``[lcc, Fisher z, relative z, x, y, z, a11 .. a33]`` for every peak;
``tofile`` writes ``solutions.out`` (:98-111).

The watershed's labelling (for each of ``steps`` cutoffs between the map's
maximum and half of it, the 6-connected components of ``corr >= cutoff`` and
each component's maximum position) runs on the GPU through
``pf_peak_positions``.  The cutoffs are computed as the reference computes
them, in the map's dtype (float32 or float64).
"""
import numpy as np

from . import _lib


def watershed_cutoffs(corr, steps):
    """The reference's cutoffs (analyzer.py:84-91), in corr's dtype."""
    max_cc = corr.max()
    min_cc = 0.5 * max_cc
    stepsize = (max_cc - min_cc) / steps
    cutoff = max_cc
    cuts = []
    for _ in range(steps):
        cutoff -= stepsize
        cuts.append(cutoff)
    return np.asarray(cuts, dtype=corr.dtype)


def peak_positions(corr, cutoffs, device=0):
    """Flat indices of the maxima of the 6-connected components of
    corr >= c for every c in cutoffs (all cutoffs, sorted, with repeats)."""
    corr = np.ascontiguousarray(corr)
    if corr.ndim != 3:
        raise ValueError("corr must be a 3-D map")
    if corr.dtype == np.float32:
        dtype = 0
    elif corr.dtype == np.float64:
        dtype = 1
    else:
        raise ValueError("corr must be float32 or float64")
    cuts = np.ascontiguousarray(cutoffs, dtype=corr.dtype)
    lib = _lib.load()
    cap = 4096
    while True:
        out = np.empty(cap, np.int64)
        count = _lib.ctypes.c_int64(0)
        _lib.check(lib.pf_peak_positions(int(device), corr.ctypes.data, dtype, *corr.shape,
                                         cuts.ctypes.data, len(cuts),
                                         out.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_int64)),
                                         cap, _lib.ctypes.byref(count)), "pf_peak_positions")
        if count.value <= cap:
            return np.sort(out[:count.value])
        cap = int(count.value)


class Analyzer:

    def __init__(self, corr, rotmat, rotmat_ind, steps=5, voxelspacing=1, origin=(0, 0, 0),
                 z_sigma=1, device=0):
        self._corr = corr
        self._rotmat = rotmat
        self._rotmat_ind = rotmat_ind
        self._voxelspacing = voxelspacing
        self._origin = origin
        self._z_sigma = z_sigma
        self._device = device
        self.steps = steps
        self._solutions = None

    @property
    def corr(self):
        return self._corr

    @property
    def rot(self):
        """The rotation-index map."""
        return self._rotmat_ind

    @property
    def steps(self):
        return self._steps

    @steps.setter
    def steps(self, steps):
        self._steps = steps
        self._watershed()
        self._solutions = None

    @property
    def voxelspacing(self):
        return self._voxelspacing

    @voxelspacing.setter
    def voxelspacing(self, voxelspacing):
        self._solutions = None
        self._voxelspacing = voxelspacing

    @property
    def origin(self):
        return self._origin

    @origin.setter
    def origin(self, origin):
        self._solutions = None
        self._origin = origin

    @property
    def solutions(self):
        if self._solutions is None:
            self._generate_solutions()
        return self._solutions

    def _solution(self, pos):
        """[lcc, Fisher z, relative z, x, y, z, R row-major] at a peak, in the
        map's scalar type as the reference computes them."""
        lcc = self._corr[pos]
        fisher = 0.5 * (np.log(1 + lcc) - np.log(1 - lcc))
        zc, yc, xc = (i * self._voxelspacing + o for i, o in zip(pos, self._origin[::-1]))
        rot = self._rotmat[int(self._rotmat_ind[pos])]
        return [lcc, fisher, fisher / self._z_sigma, xc, yc, zc, *rot.ravel()]

    def _generate_solutions(self):
        rows = [self._solution(tuple(int(i) for i in pos)) for pos in self._positions]
        self._solutions = sorted(rows, key=lambda row: row[0], reverse=True)

    def _watershed(self):
        """Positions of high correlation: per cutoff, the maximum of every
        connected region of the map above it."""
        corr = np.asarray(self._corr)
        flat = peak_positions(corr, watershed_cutoffs(corr, self._steps), self._device)
        self._positions = set(tuple(int(i) for i in p)
                              for p in zip(*np.unravel_index(flat, corr.shape)))

    def tofile(self, out="solutions.out"):
        """solutions.out: one row per solution, best first (analyzer.py:98-111)."""
        rows = self.solutions
        with open(out, "w") as fh:
            fh.write(" ".join(f"{name:{a}{w}}" for name, (a, w, _) in _COLUMNS) + "\n")
            for rank, row in enumerate(rows, start=1):
                cells = [f"{rank:<6d}"] + [f"{v:{w}.{d}f}" for v, (_, w, d)
                                           in zip(row, (c for _, c in _COLUMNS[1:]))]
                fh.write(" ".join(cells) + "\n")


# solutions.out columns: header, (header alignment, width, decimals)
_COLUMNS = ([("#rank", ("<", 6, 0))]
            + [(h, (">", 6, 3)) for h in ("cc", "Fish-z", "rel-z")]
            + [(h, (">", 8, 3)) for h in ("x", "y", "z")]
            + [(f"a{i}{j}", (">", 6, 3)) for i in (1, 2, 3) for j in (1, 2, 3)])
