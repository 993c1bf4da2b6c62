"""``PowerFitter``: the search entry point, with the reference's interface.

Mirrors ``src/powerfit_em/powerfitter.py:50-163``: construct with a target
``Volume`` (anything with ``.array``) and the Laplace flag, assign
``_rotations``, ``_template``, ``_mask`` (Volumes), optionally ``_queues``
and ``directory``, call ``scan()``, read ``_lcc`` and ``_rot``.

The reference runs forked CPU workers when ``_queues`` is None and one
OpenCL device otherwise.  Here the search always runs on a GPU through
libpowerfit_hip.so: ``_queues`` may hold device ordinals (only the first is
used, as the reference uses ``queues[0]``); None means device 0.  ``_nproc``
is accepted and ignored.  There is no CPU fallback.
"""
from os.path import abspath, isdir

import numpy as np

from .correlator import HipCorrelator
from .volume import Volume  # noqa: F401  (re-exported: PowerFitter takes Volumes)


def _array(v):
    return v.array if hasattr(v, "array") else np.asarray(v)


class PowerFitter:
    """Wrapper that runs the LCC search and exposes ``_lcc``/``_rot``."""

    def __init__(self, target, laplace=False):
        self._target = target
        self._rotations = None
        self._template = None
        self._mask = None
        self._queues = None
        self._nproc = 1
        self._directory = abspath("./")
        self._laplace = laplace
        self.mode = "auto"
        self.batch = 256

    @property
    def directory(self):
        return self._directory

    @directory.setter
    def directory(self, directory):
        if isdir(directory):
            self._directory = abspath(directory)
        else:
            raise ValueError("Directory does not exist.")

    def _device(self):
        if self._queues is None:
            return 0
        q = self._queues[0]
        return int(q) if isinstance(q, (int, np.integer)) else 0

    def scan(self):
        self._corr = HipCorrelator(_array(self._target), laplace=self._laplace,
                                   device=self._device(), mode=self.mode, batch=self.batch)
        self._corr.template = _array(self._template)
        self._corr.mask = _array(self._mask)
        self._corr.rotations = self._rotations
        self._corr.scan()
        self._lcc = self._corr.lcc
        self._rot = self._corr.rot
