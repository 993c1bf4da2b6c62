"""Split-precision FFT passes (MFMA first stage, PF_FLAG_SPLIT_DFT) end to end
against the FP64 oracle, beside the FP32 passes on the same rotations.

Bar: as every search, best LCC within 1e-4 absolute, identical rotation index
wherever the oracle's winner leads by more than 1e-4."""
import numpy as np
import pytest

from powerfit_amd import HipCorrelator
from powerfit_amd.synth import Workload

from parity import assert_scan_parity
from test_gpu_scan import case_and_oracle, say

pytestmark = pytest.mark.gpu


def run(c, R, precision, batch=8):
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="auto",
                         batch=batch, precision=precision)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    return corr.lcc, corr.rot


@pytest.mark.parametrize("name,nrot", [("t32", 48), ("t32_lap_core", 48), ("cpu64", 32),
                                       ("prec192", 6)])
def test_split_precision_vs_oracle(name, nrot):
    c, R, ref = case_and_oracle(name, nrot=nrot)
    say(f"split {name} shape={c.target.shape} first {nrot} rotations")
    best, rot = run(c, R, "split")
    err, dec = assert_scan_parity(best, rot, ref)
    b32, _ = run(c, R, "fp32")
    err32 = float(np.abs(b32 - ref.best).max())
    say(f"  max |dLCC| split {err:.2e}, fp32 {err32:.2e}; decisive voxels {dec}")


def test_split_rejects_unsupported():
    from powerfit_amd.synth import make_case
    c = make_case(Workload("odd", 0, 2.5, 200, 7.0, 20.0, seed=7300, shape=(27, 25, 30)))
    with pytest.raises(RuntimeError):
        HipCorrelator(c.target.astype(np.float64), precision="split")
    c, R, _ = case_and_oracle("t32", nrot=4)
    corr = HipCorrelator(c.target.astype(np.float64), mode="fused", precision="split")
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    with pytest.raises(RuntimeError):
        corr.scan()
