"""Rotation sets: Karney quaternion tables -> 3x3 rotation matrices.

Mirrors ``src/powerfit_em/rotations.py``:

* ``quat_to_rotmat`` (:30-61): quaternion (w, x, y, z) -> R with s = 2/|q|^2,
  rounded to 8 decimals.  The rounding is part of the contract: it fixes the
  FP64 coordinates the nearest-neighbour mask rotation rounds, and so the
  index maps.
* ``proportional_orientations`` (:88-131): pick the set whose nominal angle is
  nearest to the requested one; return (quaternions, weights, alpha).  Row
  order of the table defines the rotation index.

The tables are the public orientation sets of C. F. F. Karney (J. Mol. Graph.
Mod. 25, 595 (2007)); the ones the workloads use are shipped under
``powerfit_amd/data`` with their sha256.  Other sets are read from
``data_dir`` (e.g. an installed PowerFit's ``data`` folder) when given.
"""
import hashlib
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# name -> (number of orientations, nominal angle in degrees)
ROT_SETS = {
    "E.npy": (1, 360.0),
    "c48u1.npy": (24, 62.8),
    "c600v.npy": (60, 44.48),
    "c48n9.npy": (216, 36.47),
    "c600vc.npy": (360, 27.78),
    "c48u27.npy": (648, 20.83),
    "c48u83.npy": (1992, 16.29),
    "c48u181.npy": (4344, 12.29),
    "c48n309.npy": (7416, 9.72),
    "c48n527.npy": (12648, 8.17),
    "c48u815.npy": (19560, 7.4),
    "c48u1153.npy": (27672, 6.6),
    "c48u1201.npy": (28824, 6.48),
    "c48u1641.npy": (39384, 5.75),
    "c48u2219.npy": (53256, 5.27),
    "c48u2947.npy": (70728, 4.71),
    "c48u3733.npy": (89592, 4.37),
    "c48u4749.npy": (113976, 4.0),
    "c48u5879.npy": (141096, 3.74),
    "c48u7111.npy": (170664, 3.53),
    "c48u8649.npy": (207576, 3.26),
}

SHA256 = {
    "E.npy": "8b060680edbeea8a3df8673e428a796645d0c2d2075b4d8401f337246ea9fd7e",
    "c48u1.npy": "a296bf6ec771f04cfaf39f115172c2e89f5ccf5e0d96ef762544f4a76bfafbb0",
    "c600v.npy": "1a134cb093db944a79e58eb7032ead26d8b5f4074b24f9cc8d731eac23561b08",
    "c48n9.npy": "8f8ee7f50d3635492cde1cc303e58c686c325ce287f75c9555f889b97a67f8c5",
    "c600vc.npy": "aa853f453d6f423728cfa0bdd3ed86427dcf4f3de9046ae55141f5acd20c2aa0",
    "c48u27.npy": "5a1b3ae5ec60a1f1f1b5d894a5755641b05ca8bae360b07bec3c47f8fb3eb829",
    "c48u83.npy": "e712dc2091a485d7699bfdecced602da2cc89f9d453c7edf966759d8af3d1972",
    "c48u181.npy": "e4e768e6c620e8ecc0c8905d9438df3a8a0db8913d6177053156968e82b58617",
    "c48n309.npy": "3a1791212a87908df2a0ac1ee49d5d488fc6a59bc9e7a498bcb1e4353c2aaf46",
    "c48u2219.npy": "54d55ab0458afef161a71bc5a36c68c3a42ad6b9f14622a31512e2433021c323",
}


def quat_to_rotmat(quaternions):
    """(K, 4) quaternions (w, x, y, z) -> (K, 3, 3) rotation matrices, 8 decimals."""
    q = np.asarray(quaternions, dtype=np.float64)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    nq = w * w + x * x + y * y + z * z
    s = np.zeros(nq.shape)
    pos = nq > 0.0
    s[pos] = 2.0 / nq[pos]
    X, Y, Z = x * s, y * s, z * s
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1.0 - (y * Y + z * Z)
    R[:, 0, 1] = x * Y - w * Z
    R[:, 0, 2] = x * Z + w * Y
    R[:, 1, 0] = x * Y + w * Z
    R[:, 1, 1] = 1.0 - (x * X + z * Z)
    R[:, 1, 2] = y * Z - w * X
    R[:, 2, 0] = x * Z - w * Y
    R[:, 2, 1] = y * Z + w * X
    R[:, 2, 2] = 1.0 - (x * X + y * Y)
    np.around(R, decimals=8, out=R)
    return R


def nearest_set(angle):
    """Name of the set whose nominal angle is closest to ``angle`` (first wins ties)."""
    best, name = None, None
    for s, (_, alpha) in ROT_SETS.items():
        d = abs(angle - alpha)
        if best is None or d < best:
            best, name = d, s
    return name


def load_set(name, data_dir=None):
    path = os.path.join(data_dir or DATA_DIR, name)
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"orientation set {name} is not shipped with powerfit_amd; pass "
            f"data_dir= pointing at a folder that holds it")
    with open(path, "rb") as fh:
        blob = fh.read()
    if name in SHA256 and data_dir is None:
        if hashlib.sha256(blob).hexdigest() != SHA256[name]:
            raise ValueError(f"orientation set {name} failed its sha256 check")
    return np.load(path, allow_pickle=False)


def proportional_orientations(angle, data_dir=None):
    """(quaternions (K, 4), weights (K,), alpha) of the set nearest to ``angle``."""
    name = nearest_set(angle)
    qw = load_set(name, data_dir)
    return qw[:, :4], qw[:, -1], ROT_SETS[name][1]


def rotations_for_angle(angle, data_dir=None):
    """Rotation matrices (K, 3, 3) of the set the CLI would use for ``angle``."""
    q, _, _ = proportional_orientations(angle, data_dir)
    return quat_to_rotmat(q)
