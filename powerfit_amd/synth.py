"""Seeded synthetic fitting cases (chain template, noisy target map, mask).

PowerFit's CLI builds its inputs from a PDB and a map
(``src/powerfit_em/powerfit.py:245-267``): the structure's centroid is moved
onto the map origin, atoms are splatted as Gaussians of
sigma = resolution / (sqrt(2) * pi) with a 4-sigma cutoff weighted by atomic
number (``volume.py:213-221``, ``_powerfit.pyx:75-140``), periodically.
Without network access there are no PDB files or EM maps here, so this module
draws a synthetic chain instead and builds the same three inputs:

* chain: a CA random walk of 3.8 A steps, reflected at a confining sphere,
  with C/N/O/S side atoms 1.2-1.9 A off each CA; weights Z = 6/7/8/16;
* template: the chain with its centroid on voxel (0, 0, 0), splatted
  periodically (the correlator rotates about voxel 0);
* target: the same chain under a random rotation and translation, splatted,
  plus N(0, (noise * max)^2) noise, stored as float32;
* mask: ``template > mask_threshold * max(template)``; optionally the
  core-weighted form (``helpers.py:26-34``).

Seeds: ``seed + 0`` atoms, ``seed + 1`` pose, ``seed + 2`` noise, with
``seed = 1000 * config`` for the numbered workloads.
"""
from dataclasses import dataclass, field

import numpy as np

Z_OF = {"C": 6.0, "N": 7.0, "O": 8.0, "S": 16.0}


@dataclass
class Workload:
    """One fitting workload: grid, voxel size, chain size, resolution, options."""
    name: str
    n: int
    voxelspacing: float
    atoms: int
    resolution: float
    angle: float
    laplace: bool = False
    core_weighted: bool = False
    seed: int = 0
    noise: float = 0.05
    mask_threshold: float = 0.05
    shape: tuple = field(default=None)

    def grid_shape(self):
        return tuple(self.shape) if self.shape else (self.n, self.n, self.n)


WORKLOADS = {
    "cpu64": Workload("cpu64", 64, 2.0, 1500, 8.0, 10.0, seed=1000),
    "lc96": Workload("lc96", 96, 1.5, 5000, 6.0, 10.0, laplace=True,
                     core_weighted=True, seed=2000),
    "batch128": Workload("batch128", 128, 1.2, 12000, 5.0, 5.0, seed=3000),
    "search128": Workload("search128", 128, 1.2, 12000, 5.0, 10.0, seed=3000),
    "prec192": Workload("prec192", 192, 1.0, 25000, 4.0, 10.0, seed=4000),
    # grids without a compiled fused size (the run-time planned pipeline):
    # 120^3 and a non-cubic, radix-7 shape of the kind `extend` produces
    "grid120": Workload("grid120", 120, 1.3, 9000, 5.5, 10.0, seed=6000),
    "grid84x96x112": Workload("grid84x96x112", 84, 1.4, 6000, 6.0, 10.0, seed=6100,
                              shape=(84, 96, 112)),
    "grid70x72x98": Workload("grid70x72x98", 70, 1.6, 4000, 6.5, 10.0, seed=6200,
                             shape=(70, 72, 98)),
    # odd lengths on every axis (x = 105: two rows per x transform)
    "grid75x81x105": Workload("grid75x81x105", 75, 1.5, 5000, 6.0, 10.0, seed=6300,
                              shape=(75, 81, 105)),
}


def random_rotation(rng):
    """Uniform random rotation matrix from a normalised Gaussian quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def confining_radius(n_atoms, box_half_angstrom):
    """Sphere radius for the walk: a protein-like compact size, capped at
    0.8 of the rotation sphere so the template is never cut by it."""
    # ~13 Da per heavy atom incl. hydrogens, 1.21 A^3 per Da
    vol = 1.21 * 13.0 * n_atoms
    compact = 1.3 * (3.0 * vol / (4.0 * np.pi)) ** (1.0 / 3.0)
    return float(min(compact, 0.8 * box_half_angstrom))


def make_chain(n_atoms, radius, rng):
    """(3, n) coordinates in Angstrom (centroid at 0) and atomic numbers."""
    n_res = (n_atoms + 3) // 4
    ca = np.zeros((n_res, 3))
    pos = np.zeros(3)
    for i in range(n_res):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        new = pos + 3.8 * d
        if np.linalg.norm(new) > radius:
            new = pos - 3.8 * d
            if np.linalg.norm(new) > radius:
                new = pos * (1.0 - 3.8 / max(np.linalg.norm(pos), 3.8))
        pos = new
        ca[i] = pos
    side = rng.standard_normal((n_res, 3, 3))
    side /= np.linalg.norm(side, axis=2, keepdims=True)
    side *= rng.uniform(1.2, 1.9, (n_res, 3, 1))
    atoms = np.concatenate([ca[:, None, :], ca[:, None, :] + side], axis=1)
    atoms = atoms.reshape(-1, 3)[:n_atoms]
    kinds = rng.choice(["C", "N", "O", "S"], size=(n_res, 3), p=[0.45, 0.25, 0.27, 0.03])
    elem = np.concatenate([np.full((n_res, 1), "C"), kinds], axis=1).reshape(-1)[:n_atoms]
    weights = np.array([Z_OF[e] for e in elem])
    atoms -= atoms.mean(axis=0)
    return atoms.T.copy(), weights


def splat(xyz_vox, weights, sigma_vox, shape, chunk=4096):
    """Periodic Gaussian splat of points (3, n) in voxel units (x, y, z rows)
    onto a (nz, ny, nx) grid; every voxel within 4 sigma gets
    w * exp(-d^2 / (2 sigma^2))."""
    nz, ny, nx = shape
    out = np.zeros(nz * ny * nx)
    ext = 4.0 * sigma_vox
    c = int(np.ceil(ext)) + 1
    off = np.arange(-c, c + 1)
    oz, oy, ox = np.meshgrid(off, off, off, indexing="ij")
    oz, oy, ox = oz.ravel(), oy.ravel(), ox.ravel()
    n = xyz_vox.shape[1]
    for s in range(0, n, chunk):
        p = xyz_vox[:, s:s + chunk]
        w = weights[s:s + chunk]
        base = np.floor(p).astype(np.int64)
        ix = base[0][:, None] + ox[None, :]
        iy = base[1][:, None] + oy[None, :]
        iz = base[2][:, None] + oz[None, :]
        d2 = (ix - p[0][:, None]) ** 2 + (iy - p[1][:, None]) ** 2 + (iz - p[2][:, None]) ** 2
        keep = d2 <= ext * ext
        val = (w[:, None] * np.exp(-d2 / (2.0 * sigma_vox ** 2)))[keep]
        flat = (np.mod(iz, nz) * ny + np.mod(iy, ny)) * nx + np.mod(ix, nx)
        out += np.bincount(flat[keep], weights=val, minlength=out.size)
    return out.reshape(shape)


def core_indices(mask):
    """Core-weighted mask: the number of successive binary erosions (6-connected
    cross, zero border) each voxel survives — ``helpers.py:26-34``."""
    from scipy.ndimage import binary_erosion
    core = np.zeros(mask.shape)
    eroded = mask > 0
    while eroded.sum() > 0:
        core += eroded
        eroded = binary_erosion(eroded)
    return core


@dataclass
class Case:
    """Inputs of one fit, as the CLI hands them to the correlator."""
    target: np.ndarray      # float32 map
    template: np.ndarray    # float64, centroid on voxel 0
    mask: np.ndarray        # float64 (0/1 or core weights)
    laplace: bool
    voxelspacing: float
    resolution: float
    angle: float
    true_rotation: np.ndarray
    true_translation: np.ndarray


def make_case(w: Workload, seed=None) -> Case:
    seed = w.seed if seed is None else seed
    shape = w.grid_shape()
    rng_atoms = np.random.default_rng(seed + 0)
    rng_pose = np.random.default_rng(seed + 1)
    rng_noise = np.random.default_rng(seed + 2)
    half = min(shape) // 2 * w.voxelspacing
    xyz, weights = make_chain(w.atoms, confining_radius(w.atoms, half), rng_atoms)
    sigma_vox = w.resolution / (np.sqrt(2.0) * np.pi) / w.voxelspacing
    template = splat(xyz / w.voxelspacing, weights, sigma_vox, shape)
    rot = random_rotation(rng_pose)
    trans = rng_pose.uniform(0, 1, 3) * np.array(shape[::-1])  # voxel units, (x, y, z)
    moved = rot @ (xyz / w.voxelspacing) + trans[:, None]
    target = splat(moved, weights, sigma_vox, shape)
    target += rng_noise.normal(0.0, w.noise * target.max(), size=shape)
    target = target.astype(np.float32)
    mask = (template > w.mask_threshold * template.max()).astype(np.float64)
    if w.core_weighted:
        mask = core_indices(mask)
    return Case(target, template, mask, w.laplace, w.voxelspacing,
                w.resolution, w.angle, rot, trans)
