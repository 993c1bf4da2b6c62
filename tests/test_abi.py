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


def _occ_bytes(nz, ny, nx):
    """The brick occupancy map: one bit per 8^3 brick, ceil(n / 8) per axis."""
    nb = ((nz + 7) // 8) * ((ny + 7) // 8) * ((nx + 7) // 8)
    return 4 * ((nb + 31) // 32)


def test_ctx_bytes_model():
    """Closed form of the admission estimate: resident inputs and keys,
    rotations, the brick occupancy map, and the larger of the scan's work
    buffers and the transient buffers of pf_set_target (16 bytes per voxel)."""
    lib = _lib.load()
    n, b, K = 128, 64, 7416
    V, H = n ** 3, n * n * (n // 2 + 1)
    resident = V * 25 + 2 * H * 8 + K * 72 + 4 * (((n // 8) ** 3 + 31) // 32)
    fused = lib.pf_ctx_bytes(n, n, n, b, _lib.MODES["fused"], 0, K)
    unfused = lib.pf_ctx_bytes(n, n, n, b, _lib.MODES["batched"], 0, K)
    plain = lib.pf_ctx_bytes(n, n, n, b, _lib.MODES["plain"], 0, K)
    split = lib.pf_ctx_bytes(n, n, n, b, _lib.MODES["auto"], _lib.FLAG_SPLIT_DFT, K)
    assert fused == split == resident + 3 * b * H * 8   # 128^3: split z stage fused
    assert unfused == resident + 3 * b * (H * 8 + V * 4)
    # a fused size without the split z stage runs split precision unfused
    n3 = 96
    V3, H3 = n3 ** 3, n3 * n3 * (n3 // 2 + 1)
    res3 = V3 * 25 + 2 * H3 * 8 + K * 72 + 4 * (((n3 // 8) ** 3 + 31) // 32)
    assert lib.pf_ctx_bytes(n3, n3, n3, b, _lib.MODES["auto"], _lib.FLAG_SPLIT_DFT, K) == \
        res3 + 3 * b * (H3 * 8 + V3 * 4)
    assert plain == resident + max(3 * (H * 8 + V * 4), 16 * V)
    # a range shorter than the batch uses only as many rotations per launch
    assert lib.pf_ctx_bytes(n, n, n, b, _lib.MODES["fused"], 0, 3) == \
        V * 25 + 2 * H * 8 + 3 * 72 + 4 * (((n // 8) ** 3 + 31) // 32) + 3 * 3 * H * 8
    assert lib.pf_ctx_bytes(n, n, n, b, _lib.MODES["fused"], 0, 1) == \
        V * 25 + 2 * H * 8 + 72 + 4 * (((n // 8) ** 3 + 31) // 32) + 16 * V
    # grids without a compiled size take the run-time planned fused pipeline in
    # auto mode (spectra only) where their planes fit, else the unfused layout
    n2 = 120 + 6
    V2, H2 = n2 ** 3, n2 * n2 * (n2 // 2 + 1)
    assert lib.pf_ctx_bytes(n2, n2, n2, b, _lib.MODES["auto"], 0, 0) == \
        V2 * 25 + 2 * H2 * 8 + _occ_bytes(n2, n2, n2) + 3 * b * H2 * 8
    assert lib.pf_ctx_bytes(n2, n2, n2, b, _lib.MODES["fused_rt"], 0, 0) == \
        V2 * 25 + 2 * H2 * 8 + _occ_bytes(n2, n2, n2) + 3 * b * H2 * 8
    # odd x (two rows per x transform) within the planes' limits: spectra only
    V4, H4 = 125 ** 3, 125 * 125 * 63
    assert lib.pf_ctx_bytes(125, 125, 125, b, _lib.MODES["auto"], 0, 0) == \
        V4 * 25 + 2 * H4 * 8 + _occ_bytes(125, 125, 125) + 3 * b * H4 * 8
    # planes of more than 16,384 voxels (odd and even x): the unfused layout
    for shape in ((135, 135, 135), (200, 200, 200)):
        nz, ny, nx = shape
        V3, H3 = nz * ny * nx, nz * ny * (nx // 2 + 1)
        assert lib.pf_ctx_bytes(nz, ny, nx, b, _lib.MODES["auto"], 0, 0) == \
            V3 * 25 + 2 * H3 * 8 + _occ_bytes(nz, ny, nx) + 3 * b * (H3 * 8 + V3 * 4)
    assert lib.pf_ctx_bytes(0, 1, 1, 1, 0, 0, 0) == -1
    assert lib.pf_ctx_bytes(8, 8, 8, 1, 0, 0, -1) == -1


def test_progress_marks_reject_bad_arguments():
    lib = _lib.load()
    t = ctypes.c_uint64(0)
    assert lib.pf_ctx_mark(None, ctypes.byref(t)) == -1
    assert lib.pf_ctx_reached(None, 1) == -1


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
