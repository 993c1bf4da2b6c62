"""The C-ABI library loads without a GPU and exports what the header declares."""
import ctypes

import numpy as np

from powerfit_amd import _lib


def test_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _lib.header_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared) == set(_lib._SIGNATURES), "ctypes signatures out of sync with header"


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.pf_abi_version() == _lib.ABI_VERSION
    assert isinstance(lib.pf_last_error(), bytes)


def test_argument_errors_need_no_device():
    lib = _lib.load()
    g = np.zeros((6, 6, 6))
    out = np.zeros_like(g)
    rc = lib.pf_rotate_grid3d(_lib.ptr(g, _lib.DP), 6, 6, 6, _lib.ptr(np.eye(3), _lib.DP), 5,
                              _lib.ptr(out, _lib.DP), 6, 6, 6, 0)
    assert rc == -1 and b"radius" in lib.pf_last_error()
    assert lib.pf_ctx_create(0, 8, 8, 8, 0, None) == -1
    assert lib.pf_scan(None, 0, 1, 1, 0) == -1


def test_ctx_bytes_model():
    lib = _lib.load()
    fused = lib.pf_ctx_bytes(128, 128, 128, 64, _lib.MODES["fused"])
    unfused = lib.pf_ctx_bytes(128, 128, 128, 64, _lib.MODES["batched"])
    plain = lib.pf_ctx_bytes(128, 128, 128, 64, _lib.MODES["plain"])
    assert 0 < plain < fused < unfused
    assert lib.pf_ctx_bytes(0, 1, 1, 1, 0) == -1


def test_more_argument_errors_need_no_device():
    """Every entry point rejects bad arguments before touching a device."""
    lib = _lib.load()
    pts, w = np.zeros((3, 2)), np.ones(2)
    grid = np.zeros((4, 4, 4))
    rc = lib.pf_blur_points(0, _lib.ptr(pts, _lib.DP), _lib.ptr(w, _lib.DP), 2, 1.0, None,
                            4, 4, 4, 0)
    assert rc == -1 and b"output grid" in lib.pf_last_error()
    rc = lib.pf_dilate_points(0, _lib.ptr(pts, _lib.DP), _lib.ptr(w, _lib.DP), -1,
                              _lib.ptr(grid, _lib.DP), 4, 4, 4, 0)
    assert rc == -1 and b"atom arrays" in lib.pf_last_error()
    corr = np.zeros((4, 4, 4), np.float32)
    count = ctypes.c_int64(0)
    assert lib.pf_peak_positions(0, corr.ctypes.data, 7, 4, 4, 4, corr.ctypes.data, 1, None, 0,
                                 ctypes.byref(count)) == -1
    assert b"dtype" in lib.pf_last_error()
    assert lib.pf_peak_positions(0, corr.ctypes.data, 0, 0, 4, 4, corr.ctypes.data, 1, None, 0,
                                 ctypes.byref(count)) == -1
    assert lib.pf_peak_positions(0, corr.ctypes.data, 0, 4, 4, 4, corr.ctypes.data, 1, None, 0,
                                 None) == -1
    for rc in (lib.pf_ctx_set_schedule(None, 1), lib.pf_ctx_set_tuning(None, 1, 1),
               lib.pf_set_rotations(None, None, 3), lib.pf_scan_graph(None, 0, 1, 1, 0),
               lib.pf_get_result(None, None, None), lib.pf_rfftn_dev(None, None, 0, 4, 4, 1, None)):
        assert rc == -1
