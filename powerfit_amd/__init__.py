"""powerfit_amd — MI355X-native engine for PowerFit's exhaustive 6-D LCC search.

Host code (this package) keeps the reference's FP64 preprocessing and
interface; the per-rotation loop runs as hand-written gfx950 HIP kernels in
libpowerfit_hip.so, reached through the C-ABI in include/powerfit_hip.h.
"""
from .correlator import BaseCorrelator, HipCorrelator
from .powerfitter import PowerFitter, Volume
from .rotations import proportional_orientations, quat_to_rotmat, rotations_for_angle

__all__ = ["BaseCorrelator", "HipCorrelator", "PowerFitter", "Volume",
           "proportional_orientations", "quat_to_rotmat", "rotations_for_angle"]
