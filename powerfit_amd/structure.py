"""Atomic models: PDB/mmCIF reading, rigid-body moves, PDB writing.

Mirrors the reference's ``Structure`` (``src/powerfit_em/structure.py:146-250``)
as far as the search's callers use it: ``fromfile`` (PDB/ENT and mmCIF),
``coor`` (3 x n), ``duplicate``, ``translate``, ``rotate``, ``select`` (by
chain for ``-c``), ``atomnumber``, ``rvdw``, ``tofile``.  Records are kept in
a numpy structured array with the reference's fields, and the PDB writer
emits the reference's fixed-column layout (``structure.py:14-25,79-117``):
ATOM/HETATM lines, a TER line after each chain of ATOM records, MODEL/ENDMDL
for several models, and END.
"""
import operator
from collections import OrderedDict

import numpy as np

from .elements import lookup

_DTYPE = [("record", np.str_, 6), ("id", np.int32), ("name", np.str_, 4),
          ("alt", np.str_, 1), ("resn", np.str_, 4), ("chain", np.str_, 2),
          ("resi", np.int32), ("i", np.str_, 1), ("x", np.float64), ("y", np.float64),
          ("z", np.float64), ("q", np.float64), ("b", np.float64), ("e", np.str_, 2),
          ("charge", np.str_, 2), ("model", np.int32)]


def _element_of(name, column):
    """Element column (76-78), else the first letter of the atom name."""
    e = column.strip()
    if e:
        return e
    for ch in name:
        if ch.isalpha():
            return ch
    return ""


def read_pdb(lines):
    """ATOM/HETATM records of a PDB file as a structured array."""
    rows = []
    model = 1
    for line in lines:
        rec = line[:6]
        if rec in ("ATOM  ", "HETATM"):
            name = line[12:16].strip()
            rows.append((rec, int(line[6:11]), name, line[16], line[17:20].strip(), line[21],
                         int(line[22:26]), line[26], float(line[30:38]), float(line[38:46]),
                         float(line[46:54]), float(line[54:60]), float(line[60:66]),
                         _element_of(name, line[76:78]), line[78:80].strip(), model))
        elif rec == "MODEL ":
            model = int(line[10:14])
    return np.array(rows, dtype=_DTYPE)


def read_mmcif(lines):
    """ATOM records of an mmCIF file (``_atom_site`` loop) as a structured array."""
    keys = OrderedDict()
    words = []
    for line in lines:
        if line.startswith("_atom_site."):
            keys[line.split(".", 1)[1].strip()] = len(keys)
        elif line.startswith("ATOM"):
            words.append(line.split())
    n = len(words)
    arr = np.zeros(n, dtype=_DTYPE)

    def col(k):
        return [w[keys[k]] for w in words]
    arr["record"] = "ATOM  "
    for field, key in (("id", "id"), ("name", "label_atom_id"), ("resn", "label_comp_id"),
                       ("chain", "label_asym_id"), ("resi", "label_seq_id"),
                       ("x", "Cartn_x"), ("y", "Cartn_y"), ("z", "Cartn_z"),
                       ("q", "occupancy"), ("b", "B_iso_or_equiv"), ("e", "type_symbol"),
                       ("charge", "pdbx_formal_charge"), ("model", "pdbx_PDB_model_num")):
        if n:
            arr[field] = col(key)
    return arr


def _atom_line(r):
    name = r["name"]
    if len(r["e"]) == 1 and len(name) != 4:
        name = " " + name
    return (f"{r['record']:6s}{int(r['id']):>5d} {name:4s}{r['alt']:1s}{r['resn']:3s} "
            f"{r['chain']:1s}{int(r['resi']):>4d}{r['i']:1s}   "
            f"{r['x']:8.3f}{r['y']:8.3f}{r['z']:8.3f}{r['q']:6.2f}{r['b']:6.2f}"
            f"          {r['e']:<2s}{r['charge']:2s}\n")


def _ter_line(r):
    return (f"TER   {int(r['id']) + 1:>5d}      {r['resn']:3s} {r['chain']:1s}"
            f"{int(r['resi']):>4d}{r['i']:1s}" + " " * 53 + "\n")


def write_pdb(data, out):
    nmodels = len(set(data["model"].tolist()))
    per = len(data) // nmodels
    with open(out, "w") as fh:
        for m in range(nmodels):
            block = data[m * per:(m + 1) * per]
            if nmodels > 1:
                fh.write(f"MODEL     {m + 1:>4d}\n")
            prev = block["chain"][0] if len(block) else None
            for k, r in enumerate(block):
                if r["chain"] != prev:
                    if block["record"][k - 1] == "ATOM  ":
                        fh.write(_ter_line(block[k - 1]))
                    prev = r["chain"]
                fh.write(_atom_line(r))
            if nmodels > 1:
                fh.write("ENDMDL\n")
        fh.write("END   \n")


class Structure:

    @classmethod
    def fromfile(cls, fid):
        """Read a PDB (.pdb, .ent) or mmCIF (.cif) file (path or open file)."""
        name = getattr(fid, "name", fid)
        ext = str(name)[-3:]
        if ext not in ("pdb", "ent", "cif"):
            raise IOError("Filetype not recognized.")
        if isinstance(fid, str):
            with open(fid) as fh:
                lines = fh.readlines()
        else:
            lines = [ln.decode() if isinstance(ln, bytes) else ln for ln in fid]
        return cls(read_mmcif(lines) if ext == "cif" else read_pdb(lines))

    def __init__(self, data):
        self.data = data

    @property
    def coor(self):
        return np.asarray([self.data["x"], self.data["y"], self.data["z"]])

    @property
    def chain_list(self):
        return np.unique(self.data["chain"])

    def duplicate(self):
        return Structure(self.data.copy())

    def translate(self, trans):
        self.data["x"] += trans[0]
        self.data["y"] += trans[1]
        self.data["z"] += trans[2]

    def rotate(self, rotmat):
        xyz = np.asarray(rotmat, dtype=np.float64) @ self.coor
        self.data["x"], self.data["y"], self.data["z"] = xyz

    def select(self, identifier, values, loperator="=="):
        """Atoms whose `identifier` field compares true with any of `values`
        (all of them for '!=')."""
        ops = {"==": operator.eq, "<": operator.lt, ">": operator.gt, ">=": operator.ge,
               "<=": operator.le, "!=": operator.ne}
        if loperator not in ops:
            raise ValueError("Logic operator not recognized.")
        if isinstance(values, str) or not hasattr(values, "__iter__"):
            values = (values,)
        values = list(values)
        sel = ops[loperator](self.data[identifier], values[0])
        for v in values[1:]:
            if loperator == "!=":
                sel &= ops[loperator](self.data[identifier], v)
            else:
                sel |= ops[loperator](self.data[identifier], v)
        return Structure(self.data[sel])

    def _property(self, k):
        elements, inverse = np.unique(self.data["e"], return_inverse=True)
        return np.asarray([lookup(e)[k] for e in elements], dtype=np.float64)[inverse]

    @property
    def atomnumber(self):
        return self._property(0)

    @property
    def rvdw(self):
        return self._property(1)

    def tofile(self, out):
        write_pdb(self.data, out)
