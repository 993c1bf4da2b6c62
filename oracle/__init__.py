"""ORACLE — test infrastructure, not product code.

A CPU (FP64 NumPy + plain C) restatement of PowerFit's exhaustive 6-D
local-cross-correlation search, used only as the checker for the HIP
implementation in ``powerfit_amd``.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.

Modules
-------
``rotate``      ctypes binding of ``rotate.c`` (restates
                ``_extensions.c:7-196`` ``rotate_grid3d``) and a vectorised
                NumPy version of the same arithmetic for small grids.
``correlator``  the correlator protocol (``powerfitter.py:166-393``):
                preprocessing, target spectra, the per-rotation
                rotate -> 3 x (rfftn, conj-multiply, irfftn) -> calc_lcc ->
                take-best loop, plus top-2 tracking so that rotation-index
                parity is decidable; and the multi-process driver with the
                reference's contiguous chunking and strict-``>`` merge
                (``powerfitter.py:94-163``).
``core``        closed-form core weighting (taxicab distance to the
                background with a zero border), the quantity
                ``helpers.py:26-34`` computes by repeated erosion.

Parity status of each function (how it is pinned) is listed in the module
docstrings; every function here is pinned by a test in ``tests/``.
"""
