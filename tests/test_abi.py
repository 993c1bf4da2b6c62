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
