"""Write tests/golden/rotate_ref.npz: rotate_grid3d vectors from the reference.

The outputs come from the reference's own ``src/powerfit_em/_extensions.c``
compiled by ``make -C oracle ref`` (flags of its setup.py, including
-ffast-math) into ``oracle/_ref``.  Inputs are seeded random grids (FP64
values for trilinear, integer core-weight-like values for nearest) and
rotations from the 10-degree set (c48n309), the icosahedral c600v set (exact
half-integer coordinates) and the cube group (c48u1).

    make -C oracle ref && python tests/golden/make_rotate_fixtures.py
"""
import glob
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from powerfit_amd.rotations import load_set, quat_to_rotmat  # noqa: E402


def load_ref():
    paths = glob.glob(os.path.join(ROOT, "oracle", "_ref", "_extensions*.so"))
    if not paths:
        raise SystemExit("build the reference extension first: make -C oracle ref")
    spec = importlib.util.spec_from_file_location("_extensions", paths[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    ref = load_ref()
    rng = np.random.default_rng(20240601)
    sets = {
        "c48n309": quat_to_rotmat(load_set("c48n309.npy")[:, :4]),
        "c600v": quat_to_rotmat(load_set("c600v.npy")[:, :4]),
        "c48u1": quat_to_rotmat(load_set("c48u1.npy")[:, :4]),
    }
    shapes = [(12, 12, 12), (9, 10, 11), (8, 13, 10)]
    out = {}
    case = 0
    for shape in shapes:
        grid = rng.standard_normal(shape)
        mask = rng.integers(0, 5, shape).astype(np.float64)
        radius = min(shape) // 2
        picks = [("c48n309", int(k)) for k in rng.choice(7416, 4, replace=False)]
        picks += [("c600v", int(k)) for k in rng.choice(60, 2, replace=False)]
        picks += [("c48u1", int(k)) for k in rng.choice(24, 2, replace=False)]
        for setname, k in picks:
            R = sets[setname][k]
            for nearest, g in ((0, grid), (1, mask)):
                res = np.zeros(shape)
                ref.rotate_grid3d(g, R, radius, res, nearest)
                out[f"case{case}_grid"] = g
                out[f"case{case}_rot"] = R
                out[f"case{case}_radius"] = np.array(radius)
                out[f"case{case}_nearest"] = np.array(nearest)
                out[f"case{case}_out"] = res
                case += 1
    out["ncases"] = np.array(case)
    path = os.path.join(HERE, "rotate_ref.npz")
    np.savez_compressed(path, **out)
    print(path, case, "cases", os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
