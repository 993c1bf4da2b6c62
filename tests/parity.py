"""Shared parity checks: GPU results against the FP64 oracle."""
import numpy as np

LCC_TOL = 1e-4   # best-LCC absolute tolerance; also the decisive-margin threshold


def scan_parity(best, rot, ref, tol=LCC_TOL):
    """Return (max |best - ref.best|, #index mismatches where margin > tol, #decisive)."""
    err = float(np.abs(best.astype(np.float64) - ref.best).max())
    decisive = ref.margin() > tol
    mism = int(((rot.astype(np.int64) != ref.rot.astype(np.int64)) & decisive).sum())
    return err, mism, int(decisive.sum())


def assert_scan_parity(best, rot, ref, tol=LCC_TOL):
    err, mism, dec = scan_parity(best, rot, ref, tol)
    assert err <= tol, f"best-LCC error {err:.3e} > {tol}"
    assert mism == 0, f"{mism} rotation-index mismatches among {dec} decisive voxels"
    return err, dec
