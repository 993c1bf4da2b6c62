"""ctypes binding of libpowerfit_hip.so (C-ABI in include/powerfit_hip.h).

The library is built in-tree by ``python -m powerfit_amd.build``.  There is no
CPU fallback: a missing library or a failing call raises.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpowerfit_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "powerfit_hip.h")

ABI_VERSION = 2
MODES = {"plain": 0, "batched": 1, "fused": 2, "auto": 3,
         "mix_fwd": 4, "mix_z": 5, "mix_inv": 6, "fused_rt": 7}
MODE_PLAIN = MODES["plain"]
MODE_BATCHED = MODES["batched"]
FLAG_NO_BINARY_MASK = 1
FLAG_SPLIT_DFT = 2
KERNELS = ["rotate", "r2c_x", "c2c_fwd", "conj_mul", "c2c_inv", "c2r_x", "lcc",
           "fwd_plane", "zcorr", "inv_plane"]

_c_int, _c_i64, _c_u32, _c_dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_uint32, ctypes.c_double
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

_SIGNATURES = {
    "pf_abi_version": ([], _c_int),
    "pf_last_error": ([], ctypes.c_char_p),
    "pf_device_count": ([ctypes.POINTER(_c_int)], _c_int),
    "pf_rotate_grid3d": ([_dp, _c_int, _c_int, _c_int, _dp, _c_int, _dp, _c_int, _c_int,
                          _c_int, _c_int], _c_int),
    "pf_conj_multiply_dev": ([_vp, _vp, _vp, _c_i64, _vp], _c_int),
    "pf_calc_lcc_dev": ([_vp, _vp, _vp, _vp, _vp, _c_i64, _vp], _c_int),
    "pf_rfftn_dev": ([_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp], _c_int),
    "pf_irfftn_dev": ([_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp], _c_int),
    "pf_ctx_create": ([_c_int, _c_int, _c_int, _c_int, _c_u32, ctypes.POINTER(_vp)], _c_int),
    "pf_ctx_destroy": ([_vp], _c_int),
    "pf_ctx_set_stream": ([_vp, _vp], _c_int),
    "pf_set_target": ([_vp, _dp, _u8p], _c_int),
    "pf_set_template_mask": ([_vp, _dp, _dp, _c_dbl, _c_int], _c_int),
    "pf_set_rotations": ([_vp, _dp, _c_i64], _c_int),
    "pf_reset_result": ([_vp], _c_int),
    "pf_scan": ([_vp, _c_i64, _c_i64, _c_int, _c_int], _c_int),
    "pf_scan_graph": ([_vp, _c_i64, _c_i64, _c_int, _c_int], _c_int),
    "pf_scan_noalloc": ([_vp, _c_i64, _c_i64, _c_int, _c_int], _c_int),
    "pf_get_result": ([_vp, _fp, _i32p], _c_int),
    "pf_sync": ([_vp], _c_int),
    "pf_ctx_profile": ([_vp, _c_int], _c_int),
    "pf_ctx_fused_available": ([_vp], _c_int),
    "pf_ctx_set_tuning": ([_vp, _c_int, _c_int], _c_int),
    "pf_ctx_query": ([_vp], _c_int),
    "pf_ctx_bytes": ([_c_int, _c_int, _c_int, _c_int, _c_int, _c_u32, _c_i64], _c_i64),
    "pf_ctx_mark": ([_vp, ctypes.POINTER(ctypes.c_uint64)], _c_int),
    "pf_ctx_reached": ([_vp, ctypes.c_uint64], _c_int),
    "pf_ctx_kernel_time": ([_vp, _c_int, _dp, ctypes.POINTER(_c_i64)], _c_int),
    "pf_ctx_set_schedule": ([_vp, _c_int], _c_int),
    "pf_blur_points": ([_c_int, _dp, _dp, _c_i64, _c_dbl, _dp, _c_int, _c_int, _c_int, _c_int],
                       _c_int),
    "pf_dilate_points": ([_c_int, _dp, _dp, _c_i64, _dp, _c_int, _c_int, _c_int, _c_int], _c_int),
    "pf_peak_positions": ([_c_int, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int,
                           ctypes.POINTER(_c_i64), _c_i64, ctypes.POINTER(_c_i64)], _c_int),
}

_lib = None


class PowerFitHipError(RuntimeError):
    """A C-ABI call returned an error code."""


def header_symbols():
    """Function names declared in include/powerfit_hip.h."""
    with open(HEADER) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_\w+)\s*\(", text)))


def load():
    """Load the library (once) and declare every entry point's signature."""
    global _lib
    if _lib is not None:
        return _lib
    # torch ships its own libamdhip64 (same SONAME as ROCm's).  Loading torch
    # first lets the dynamic loader resolve this library's HIP dependency to
    # that copy, so a process that uses both has a single HIP runtime; loading
    # the system copy first would leave torch unable to see the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m powerfit_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.pf_abi_version() != ABI_VERSION:
        raise ImportError("libpowerfit_hip.so ABI version mismatch")
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = load().pf_last_error().decode(errors="replace")
        raise PowerFitHipError(f"{what or 'call'} failed ({rc}): {msg}")
    return rc


def ptr(a, kind):
    """ctypes pointer to a contiguous numpy array of the right dtype."""
    return a.ctypes.data_as(kind)


DP, FP, I32P, U8P = _dp, _fp, _i32p, _u8p
