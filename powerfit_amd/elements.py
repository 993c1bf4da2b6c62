"""Atomic numbers and van der Waals radii (A) of the elements that occur in
biomolecular models.

``Structure.atomnumber`` weights the template density by atomic number
(``structure.py:167-170`` of the reference, used at ``powerfit.py:245-252``)
and ``Structure.rvdw`` gives the radii of the molecular-volume mask
(``powerfit.py:286-296``).  The radii are Bondi's (J. Phys. Chem. 68, 441,
1964) as listed in the reference's element table
(``src/powerfit_em/elements.py``), 0.0 where that table has none.
"""

# symbol: (atomic number, van der Waals radius)
TABLE = {
    "H": (1, 1.2), "He": (2, 1.4), "Li": (3, 1.82), "B": (5, 0.0), "C": (6, 1.7),
    "N": (7, 1.55), "O": (8, 1.52), "F": (9, 1.47), "Ne": (10, 1.54), "Na": (11, 2.27),
    "Mg": (12, 1.73), "Al": (13, 0.0), "Si": (14, 2.1), "P": (15, 1.8), "S": (16, 1.8),
    "Cl": (17, 1.75), "Ar": (18, 1.88), "K": (19, 2.75), "Ca": (20, 0.0), "V": (23, 0.0),
    "Cr": (24, 0.0), "Mn": (25, 0.0), "Fe": (26, 0.0), "Co": (27, 0.0), "Ni": (28, 1.63),
    "Cu": (29, 1.4), "Zn": (30, 1.39), "Se": (34, 1.9), "Br": (35, 1.85), "Kr": (36, 2.02),
    "Rb": (37, 0.0), "Sr": (38, 0.0), "Mo": (42, 0.0), "Ag": (47, 1.72), "Cd": (48, 1.58),
    "I": (53, 1.98), "Xe": (54, 2.16), "Cs": (55, 0.0), "Ba": (56, 0.0), "W": (74, 0.0),
    "Pt": (78, 1.75), "Au": (79, 1.66), "Hg": (80, 0.0),
}


def lookup(symbol):
    """(atomic number, vdW radius) of an element symbol in any letter case."""
    key = symbol.strip().capitalize()
    if key not in TABLE:
        raise KeyError(f"element {symbol!r} is not in the table")
    return TABLE[key]
