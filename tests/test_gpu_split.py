"""Split-precision FFT stages (MFMA first stage, PF_FLAG_SPLIT_DFT) end to end
against the FP64 oracle, beside the FP32 path on the same rotations: the
unfused y and z passes, and the fused z stage at 128^3.

Bar: as every search, best LCC within 1e-4 absolute, identical rotation index
wherever the oracle's winner leads by more than 1e-4."""
import numpy as np
import pytest

from powerfit_amd import HipCorrelator
from powerfit_amd.synth import Workload

from parity import assert_scan_parity
from test_gpu_scan import case_and_oracle, say

pytestmark = pytest.mark.gpu


def run(c, R, precision, batch=8, mode="auto", **kw):
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode=mode,
                         batch=batch, precision=precision, **kw)
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


@pytest.mark.parametrize("mode,kw", [("fused", {}), ("fused", {"binary_shortcut": False}),
                                     ("mix_z", {})])
def test_split_fused_z_stage_128_vs_oracle(mode, kw):
    """The fused z stage's split first stage at 128^3 (0/1 mask, the W^2 grid
    as its own transform, and alone beside unfused split passes): oracle
    parity, and within 1e-5 of the FP32 fused path."""
    c, R, ref = case_and_oracle("search128", nrot=8)
    say(f"split fused z stage 128^3, mode {mode} {kw}, 8 rotations")
    best, rot = run(c, R, "split", mode=mode, **kw)
    err, dec = assert_scan_parity(best, rot, ref)
    b32, _ = run(c, R, "fp32", mode="fused", **kw)
    d = float(np.abs(best - b32).max())
    say(f"  max |dLCC| split {err:.2e}; |split - fp32 fused| {d:.2e}; decisive voxels {dec}")
    assert d <= 1e-5


def test_split_fused_core_weighted_128_vs_oracle():
    """Core-weighted mask at 128^3 (three forward grids) with the split z stage."""
    from powerfit_amd.rotations import rotations_for_angle
    from powerfit_amd.synth import make_case
    from oracle import correlator as oc
    from test_gpu_scan import NPROC
    c = make_case(Workload("w128s", 128, 1.2, 6000, 5.0, 10.0, core_weighted=True, seed=3021))
    R = rotations_for_angle(10.0)[:6]
    say("split fused z stage 128^3 core-weighted, 6 rotations")
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    best, rot = run(c, R, "split", mode="fused")
    assert_scan_parity(best, rot, ref)


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
    # a compiled fused size whose z plan has no radix-16 first stage
    c, R, _ = case_and_oracle("cpu64", nrot=4)
    corr = HipCorrelator(c.target.astype(np.float64), mode="fused", precision="split")
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    with pytest.raises(RuntimeError):
        corr.scan()
