"""Write golden vectors for the output side of the fitting workflow from the
reference's own Python, imported in this container from its source files:

* analyzer_ref.npz / analyzer_ref_solutions.out -- ``Analyzer``
  (src/powerfit_em/analyzer.py): the watershed positions, the solution rows
  and the ``solutions.out`` text for a seeded LCC map;
* structure_ref_*.pdb -- ``Structure`` (src/powerfit_em/structure.py):
  a PDB read and written back, and the centred, rotated and translated copy
  that ``write_fits_to_pdb`` (helpers.py:54-64) writes for a solution row;
  plus atomic numbers and van der Waals radii per atom.

The input PDB (structure_in.pdb) is written by this script: two chains of
ATOM records with a TER between them, HETATM records, four-letter and
two-letter element names, and a line without the element column.

    python tests/golden/make_output_fixtures.py

The reference modules are loaded from /root/reference without its package
``__init__`` (which needs the compiled Cython extension); analyzer.py needs
only numpy and scipy, structure.py its sibling elements.py.
"""
import importlib.util
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/powerfit_em"


def load(name):
    full = f"powerfit_em.{name}"
    spec = importlib.util.spec_from_file_location(full, os.path.join(SRC, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[full] = mod
    spec.loader.exec_module(mod)
    return mod


PDB_IN = """\
ATOM      1  N   MET A   1      11.104  13.207   2.100  1.00 20.00           N
ATOM      2  CA  MET A   1      12.560  13.380   2.300  1.00 21.50           C
ATOM      3  C   MET A   1      13.100  14.800   2.050  0.50 22.00           C
ATOM      4  O   MET A   1      12.400  15.790   2.210  1.00 23.25           O
ATOM      5  SD  MET A   1      14.020  11.900   4.400  1.00 30.00           S
ATOM      6 HD21 ASN A   2      10.500  16.210   1.000  1.00 10.00           H
ATOM      7  N   GLY B   5      -3.250   4.500  -7.125  1.00 15.00           N
ATOM      8  CA AGLY B   5      -4.000   5.750  -8.000  0.60 16.00           C
ATOM      9  CA BGLY B   5      -4.050   5.700  -8.050  0.40 16.00           C
ATOM     10  C   GLY B   5A     -5.500   5.250  -8.500  1.00 17.00
HETATM   11 ZN   ZN  B 101       0.000   0.000   0.000  1.00 40.00          ZN2+
HETATM   12  O   HOH C 201       1.500  -2.500   3.000  1.00 35.00           O
END
"""


def main():
    analyzer = load("analyzer")
    load("elements")
    structure = load("structure")

    # Analyzer on a smooth seeded map with distinct values
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(20250)
    corr = gaussian_filter(rng.standard_normal((12, 14, 16)), 1.2, mode="wrap")
    corr = (0.9 * corr / np.abs(corr).max()).astype(np.float32)
    rotmat = rng.standard_normal((7, 3, 3))
    rot_ind = rng.integers(0, 7, corr.shape).astype(np.int32)
    params = dict(steps=5, voxelspacing=1.7, origin=(3.5, -2.25, 10.0), z_sigma=0.37)
    a = analyzer.Analyzer(corr, rotmat, rot_ind, **params)
    positions = np.array(sorted(a._positions), dtype=np.int64)
    sols = np.array([[float(v) for v in s] for s in a.solutions])
    out = os.path.join(HERE, "analyzer_ref_solutions.out")
    a.tofile(out)
    np.savez_compressed(os.path.join(HERE, "analyzer_ref.npz"), corr=corr, rotmat=rotmat,
                        rot_ind=rot_ind, steps=params["steps"],
                        voxelspacing=params["voxelspacing"], origin=np.array(params["origin"]),
                        z_sigma=params["z_sigma"], positions=positions, solutions=sols)

    # Structure: read, write back, and one write_fits_to_pdb model
    pin = os.path.join(HERE, "structure_in.pdb")
    with open(pin, "w") as fh:
        fh.write(PDB_IN)
    s = structure.Structure.fromfile(pin)
    s.tofile(os.path.join(HERE, "structure_ref_roundtrip.pdb"))
    np.savez_compressed(os.path.join(HERE, "structure_ref.npz"), coor=s.coor,
                        atomnumber=s.atomnumber, rvdw=s.rvdw)
    # write_fits_to_pdb's steps (helpers.py:54-64) on the reference Structure
    sol = [0.8, 1.1, 2.97, 31.5, -4.25, 17.0] + list(
        np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]).ravel())
    t = s.duplicate()
    t.translate(-t.coor.mean(axis=1))
    m = t.duplicate()
    m.rotate(np.asarray([float(x) for x in sol[6:]]).reshape(3, 3))
    m.translate(sol[3:6])
    m.tofile(os.path.join(HERE, "structure_ref_fit_1.pdb"))
    np.save(os.path.join(HERE, "structure_ref_solution.npy"), np.array(sol))
    print("wrote analyzer and structure fixtures")


if __name__ == "__main__":
    main()
