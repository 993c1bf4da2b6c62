"""Correlators with PowerFit's correlator protocol, backed by the HIP engine.

``BaseCorrelator`` restates the reference's host-side state machine
(``src/powerfit_em/powerfitter.py:166-255``): target normalisation and lcc
mask, template/mask/rotations setters with the same ValueErrors, Laplace
filtering, and template normalisation over the mask — all FP64 NumPy, as in
the reference.  ``HipCorrelator`` replaces ``CPUCorrelator``/``GPUCorrelator``
(``:258-556``): the per-rotation loop (rotate -> 3 x (R2C, conj-multiply,
C2R) -> calc_lcc -> take-best) runs in libpowerfit_hip.so on one GPU, and
the results ``lcc`` (float32) and ``rot`` (int32) have the shape of the
target, like ``GPUCorrelator``'s.
"""
import ctypes

import numpy as np
from scipy.ndimage import laplace

from . import _lib


class BaseCorrelator:
    """Target/template/mask/rotation state shared by all correlators."""

    def __init__(self, target, laplace=False):
        target = np.asarray(target, dtype=np.float64)
        self._target = target / target.max()
        self._rotations = None
        self._template = None
        self._mask = None
        self._laplace = laplace
        self._lcc_mask = self._get_lcc_mask(self._target)
        self._rmax = min(target.shape) // 2

    @staticmethod
    def _get_lcc_mask(target):
        return (target > target.max() * 0.05).astype(np.uint8)

    @property
    def target(self):
        return self._target

    @property
    def mask(self):
        return self._mask

    @mask.setter
    def mask(self, mask):
        if self._template is None:
            raise ValueError("First set the template.")
        mask = np.asarray(mask, dtype=np.float64) if mask is not None else mask
        if mask is None or self._target.shape != mask.shape:
            raise ValueError("Shape of the mask is different from target.")
        ind = mask != 0
        self._norm_factor = ind.sum()
        if self._norm_factor == 0:
            raise ValueError("Zero-filled mask is not allowed.")
        self._mask = mask.copy()
        if self._laplace:
            self._template = self._laplace_filter(self._template)
        self._template *= self._mask
        self._normalize_template(ind)
        # multiply again for the core-weighted correlation score
        self._template *= self._mask

    @staticmethod
    def _laplace_filter(array):
        return laplace(array, mode="wrap")

    def _normalize_template(self, ind):
        self._template[ind] -= self._template[ind].mean()
        self._template[ind] /= self._template[ind].std()

    @property
    def rotations(self):
        return self._rotations

    @rotations.setter
    def rotations(self, rotations):
        self._rotations = np.asarray(rotations, dtype=np.float64).reshape(-1, 3, 3)

    @property
    def template(self):
        return self._template

    @template.setter
    def template(self, template):
        template = np.asarray(template, dtype=np.float64)
        if template.shape != self._target.shape:
            raise ValueError("Shape of template does not match the target.")
        self._mask = None
        self._template = template.copy()

    @property
    def lcc(self):
        return self._lcc

    @property
    def rot(self):
        return self._rot

    def scan(self):
        if any(req is None for req in (self._template, self._mask, self._rotations)):
            raise ValueError("First set the template, mask, and rotations.")


class HipCorrelator(BaseCorrelator):
    """LCC search on one MI355X through libpowerfit_hip.so.

    mode  : "plain" (one rotation per launch chain, unfused kernels),
            "batched" (``batch`` rotations per launch, unfused kernels),
            "fused" (``batch`` rotations per launch through the three fused
            kernels; cubic grids of a compiled size), "fused_rt" (the three
            fused stages with the grid shape taken at run time: other grids
            whose planes fit LDS, at most 16,384 voxels each), or "auto"
            (fused, else fused_rt where
            available, else batched).  "mix_fwd"/"mix_z"/"mix_inv" run one
            fused stage with the rest unfused (for testing).
    batch : rotations per launch (256 measured best at every fused size; a
        scan of fewer rotations uses only as many).
    binary_shortcut : use FFT(W^2) = FFT(W) when the mask is 0/1.
    precision : "fp32" (default) or "split": the first radix-16 stage of an
        FFT runs as a DFT-as-GEMM on the matrix cores with FP16 hi + lo
        operands and FP32 accumulation -- in the unfused y and z passes (ny, nz
        divisible by 16 and at most 320), and in the fused pipeline's forward z
        transforms at the compiled sizes whose z plan starts with radix 16
        (128^3); other fused sizes and the run-time planned pipeline are not
        used with it.
    """

    def __init__(self, target, laplace=False, device=0, mode="auto", batch=256,
                 binary_shortcut=True, precision="fp32"):
        super().__init__(target, laplace=laplace)
        if mode not in _lib.MODES:
            raise ValueError(f"unknown mode {mode!r}")
        if precision not in ("fp32", "split"):
            raise ValueError(f"unknown precision {precision!r}")
        self.precision = precision
        self._dev = device
        self.mode = mode
        self.batch = int(batch)
        # the library's default for this grid (pf_ctx_create): the register
        # z stage at 64^3, 72^3, 96^3, 160^3 and 192^3, else schedule 0
        n = self._target.shape
        self.z_schedule = 4 if n[0] == n[1] == n[2] and n[0] in (64, 72, 96, 160, 192) else 0
        lib = _lib.load()
        self._libref = lib
        ctx = ctypes.c_void_p()
        nz, ny, nx = self._target.shape
        flags = 0 if binary_shortcut else _lib.FLAG_NO_BINARY_MASK
        if precision == "split":
            flags |= _lib.FLAG_SPLIT_DFT
        _lib.check(lib.pf_ctx_create(device, nz, ny, nx, flags, ctypes.byref(ctx)),
                   "pf_ctx_create")
        self._ctx = ctx
        tf = self._laplace_filter(self._target) if self._laplace else self._target
        tf = np.ascontiguousarray(tf, dtype=np.float64)
        mask = np.ascontiguousarray(self._lcc_mask, dtype=np.uint8)
        _lib.check(lib.pf_set_target(ctx, _lib.ptr(tf, _lib.DP), _lib.ptr(mask, _lib.U8P)),
                   "pf_set_target")

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            self._libref.pf_ctx_destroy(ctx)
            self._ctx = None

    @property
    def mask(self):
        return self._mask

    @mask.setter
    def mask(self, mask):
        BaseCorrelator.mask.fset(self, mask)
        t = np.ascontiguousarray(self._template, dtype=np.float64)
        m = np.ascontiguousarray(self._mask, dtype=np.float64)
        _lib.check(self._libref.pf_set_template_mask(
            self._ctx, _lib.ptr(t, _lib.DP), _lib.ptr(m, _lib.DP),
            float(self._norm_factor), int(self._rmax)), "pf_set_template_mask")

    @property
    def rotations(self):
        return self._rotations

    @rotations.setter
    def rotations(self, rotations):
        BaseCorrelator.rotations.fset(self, rotations)
        r = np.ascontiguousarray(self._rotations, dtype=np.float64)
        _lib.check(self._libref.pf_set_rotations(self._ctx, _lib.ptr(r, _lib.DP), r.shape[0]),
                   "pf_set_rotations")

    def scan(self, first=0, count=None):
        super().scan()
        nrot = self._rotations.shape[0]
        count = nrot - first if count is None else count
        mode = _lib.MODES[self.mode]
        lib = self._libref
        _lib.check(lib.pf_reset_result(self._ctx), "pf_reset_result")
        _lib.check(lib.pf_scan(self._ctx, first, count, self.batch, mode), "pf_scan")
        self._lcc = np.zeros(self._target.shape, dtype=np.float32)
        self._rot = np.zeros(self._target.shape, dtype=np.int32)
        _lib.check(lib.pf_get_result(self._ctx, _lib.ptr(self._lcc, _lib.FP),
                                     _lib.ptr(self._rot, _lib.I32P)), "pf_get_result")

    # -- measurement helpers -------------------------------------------------
    def set_stream(self, stream_handle):
        _lib.check(self._libref.pf_ctx_set_stream(self._ctx, ctypes.c_void_p(stream_handle)),
                   "pf_ctx_set_stream")

    def profile(self, enable=True):
        _lib.check(self._libref.pf_ctx_profile(self._ctx, int(enable)), "pf_ctx_profile")

    def kernel_times(self):
        """{kernel class: (total ms, launches)} since profiling was enabled."""
        out = {}
        for i, name in enumerate(_lib.KERNELS):
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            _lib.check(self._libref.pf_ctx_kernel_time(self._ctx, i, ctypes.byref(ms),
                                                       ctypes.byref(n)), "pf_ctx_kernel_time")
            out[name] = (ms.value, n.value)
        return out

    def enqueue_scan(self, first, count):
        """Enqueue a scan of [first, first+count) without reset or sync."""
        _lib.check(self._libref.pf_scan(self._ctx, first, count, self.batch,
                                        _lib.MODES[self.mode]), "pf_scan")

    def set_tuning(self, chunk_z=0, chunk_lcc=0):
        """Rotations per block in the fused z and LCC kernels (0 = automatic)."""
        _lib.check(self._libref.pf_ctx_set_tuning(self._ctx, int(chunk_z), int(chunk_lcc)),
                   "pf_ctx_set_tuning")

    def set_schedule(self, z_schedule=0):
        """Schedule of the fused z stage (0 .. 5; bit 0 the alternative radix
        plan, bit 1 the alternative column tile width, 4 and 5 the run-time
        planned pipeline's register z stage with its default or alternative
        register plan; 0 is the default).  The run-time planned z stage reads
        bit 0 only (its alternative register plan, where the z length has
        one)."""
        _lib.check(self._libref.pf_ctx_set_schedule(self._ctx, int(z_schedule)),
                   "pf_ctx_set_schedule")
        self.z_schedule = int(z_schedule)

    @property
    def fused_available(self):
        """The compiled fused pipeline exists for this grid."""
        return self._libref.pf_ctx_fused_available(self._ctx) == 1

    @property
    def fused_rt_available(self):
        """Only the run-time planned fused pipeline exists for this grid."""
        return self._libref.pf_ctx_fused_available(self._ctx) == 2

    def enqueue_scan_graph(self, first, count):
        """Enqueue [first, first+count) as one replay of a captured HIP graph."""
        _lib.check(self._libref.pf_scan_graph(self._ctx, first, count, self.batch,
                                              _lib.MODES[self.mode]), "pf_scan_graph")

    def done(self):
        """True when all enqueued work on this correlator's stream has finished."""
        rc = self._libref.pf_ctx_query(self._ctx)
        if rc < 0:
            _lib.check(rc, "pf_ctx_query")
        return rc == 1

    def mark(self):
        """Progress ticket for the work enqueued so far (see ``reached``)."""
        t = ctypes.c_uint64(0)
        _lib.check(self._libref.pf_ctx_mark(self._ctx, ctypes.byref(t)), "pf_ctx_mark")
        return t.value

    def reached(self, ticket):
        """True once the work enqueued before ``mark()`` returned `ticket` has
        finished (does not block)."""
        rc = self._libref.pf_ctx_reached(self._ctx, int(ticket))
        if rc < 0:
            _lib.check(rc, "pf_ctx_reached")
        return rc == 1

    def result(self):
        """Copy out (lcc float32, rot int32) of the work enqueued so far (syncs)."""
        self._lcc = np.zeros(self._target.shape, dtype=np.float32)
        self._rot = np.zeros(self._target.shape, dtype=np.int32)
        _lib.check(self._libref.pf_get_result(self._ctx, _lib.ptr(self._lcc, _lib.FP),
                                              _lib.ptr(self._rot, _lib.I32P)), "pf_get_result")
        return self._lcc, self._rot

    def reset(self):
        _lib.check(self._libref.pf_reset_result(self._ctx), "pf_reset_result")

    def sync(self):
        _lib.check(self._libref.pf_sync(self._ctx), "pf_sync")
