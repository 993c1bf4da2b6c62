"""Output formats of the fitting workflow against golden files written by the
reference's own code (tests/golden/make_output_fixtures.py): the solution rows
and ``solutions.out`` of ``Analyzer`` (analyzer.py:56-111), and the PDB files
of ``Structure.tofile`` and ``write_fits_to_pdb`` (structure.py:79-117,
helpers.py:54-64).  CPU only: the watershed positions come from the fixture.
"""
import os

import numpy as np

from powerfit_amd.analyzer import Analyzer
from powerfit_amd.helpers import fisher_sigma, write_fits_to_pdb
from powerfit_amd.structure import Structure

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _analyzer_from_fixture():
    f = np.load(os.path.join(GOLD, "analyzer_ref.npz"))
    a = Analyzer.__new__(Analyzer)          # positions from the fixture, no GPU watershed
    a._corr, a._rotmat, a._rotmat_ind = f["corr"], f["rotmat"], f["rot_ind"]
    a._voxelspacing = float(f["voxelspacing"])
    a._origin = tuple(float(v) for v in f["origin"])
    a._z_sigma = float(f["z_sigma"])
    a._steps = int(f["steps"])
    a._positions = set(tuple(int(v) for v in p) for p in f["positions"])
    a._solutions = None
    return a, f


def test_solution_rows_equal_reference():
    a, f = _analyzer_from_fixture()
    got = np.array([[float(v) for v in row] for row in a.solutions])
    assert got.shape == f["solutions"].shape and got.shape[0] > 3
    assert np.array_equal(got, f["solutions"])


def test_solutions_file_equals_reference(tmp_path):
    a, _ = _analyzer_from_fixture()
    out = tmp_path / "solutions.out"
    a.tofile(str(out))
    with open(os.path.join(GOLD, "analyzer_ref_solutions.out")) as fh:
        assert out.read_text() == fh.read()


def test_pdb_round_trip_equals_reference(tmp_path):
    s = Structure.fromfile(os.path.join(GOLD, "structure_in.pdb"))
    out = tmp_path / "rt.pdb"
    s.tofile(str(out))
    with open(os.path.join(GOLD, "structure_ref_roundtrip.pdb")) as fh:
        assert out.read_text() == fh.read()
    f = np.load(os.path.join(GOLD, "structure_ref.npz"))
    assert np.array_equal(s.coor, f["coor"])
    assert np.array_equal(s.atomnumber, f["atomnumber"])
    assert np.array_equal(s.rvdw, f["rvdw"])


def test_write_fits_to_pdb_equals_reference(tmp_path):
    s = Structure.fromfile(os.path.join(GOLD, "structure_in.pdb"))
    sol = np.load(os.path.join(GOLD, "structure_ref_solution.npy"))
    write_fits_to_pdb(s, [list(sol)], basename=str(tmp_path / "fit"))
    with open(os.path.join(GOLD, "structure_ref_fit_1.pdb")) as fh:
        assert (tmp_path / "fit_1.pdb").read_text() == fh.read()
    # the input structure is not moved
    assert np.array_equal(s.coor, Structure.fromfile(os.path.join(GOLD, "structure_in.pdb")).coor)


def test_structure_moves_and_selection():
    s = Structure.fromfile(os.path.join(GOLD, "structure_in.pdb"))
    R = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    t = s.duplicate()
    t.rotate(R)
    t.translate([1.0, 2.0, 3.0])
    assert np.allclose(t.coor, R @ s.coor + np.array([[1.0], [2.0], [3.0]]))
    assert list(s.select("chain", ["A", "C"]).data["id"]) == [1, 2, 3, 4, 5, 6, 12]
    assert list(s.select("chain", "B").data["id"]) == [7, 8, 9, 10, 11]
    assert list(s.chain_list) == ["A", "B", "C"]


def test_mmcif_reader(tmp_path):
    cif = tmp_path / "m.cif"
    cif.write_text(
        "data_x\nloop_\n_atom_site.group_PDB\n_atom_site.id\n_atom_site.type_symbol\n"
        "_atom_site.label_atom_id\n_atom_site.label_comp_id\n_atom_site.label_asym_id\n"
        "_atom_site.label_seq_id\n_atom_site.Cartn_x\n_atom_site.Cartn_y\n_atom_site.Cartn_z\n"
        "_atom_site.occupancy\n_atom_site.B_iso_or_equiv\n_atom_site.pdbx_formal_charge\n"
        "_atom_site.pdbx_PDB_model_num\n"
        "ATOM 1 N N MET A 1 1.0 2.0 3.0 1.00 20.0 ? 1\n"
        "ATOM 2 C CA MET A 1 2.5 3.5 4.5 1.00 21.0 ? 1\n")
    s = Structure.fromfile(str(cif))
    assert np.array_equal(s.coor, [[1.0, 2.5], [2.0, 3.5], [3.0, 4.5]])
    assert list(s.atomnumber) == [7.0, 6.0]


def test_fisher_sigma_closed_form():
    assert fisher_sigma(1003.0, 10.0) == 1 / np.sqrt(1003.0 / 10.0 - 3)
