"""GPU parity of the forms and rules the benchmark workloads do not reach:
weighted masks through the fused kernels, W^2 transformed separately, the
strict-'>' tie rule, NaN from a zero variance, graphs after the inputs change,
and the plain path at the 96^3 configuration.

Bar: best LCC within 1e-4 of the FP64 oracle; rotation indices identical
where the oracle's winner leads by more than 1e-4, and bit-identical maps
wherever a rule fixes the result exactly (ties, 0/0).
"""
import os
import sys

import numpy as np
import pytest

from oracle import correlator as oc
from powerfit_amd import HipCorrelator
from powerfit_amd.rotations import rotations_for_angle
from powerfit_amd.synth import WORKLOADS, Workload, make_case

from parity import assert_scan_parity
from special_cases import cube_group, zero_variance_case

pytestmark = pytest.mark.gpu
NPROC = min(8, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(target, template, mask, R, laplace=False, mode="fused", batch=8, **kw):
    corr = HipCorrelator(np.asarray(target, np.float64), laplace=laplace, mode=mode, batch=batch,
                         **kw)
    corr.template = template
    corr.mask = mask
    corr.rotations = R
    corr.scan()
    return corr.lcc.copy(), corr.rot.copy()


WEIGHTED = {
    48: Workload("w48", 48, 2.0, 600, 8.0, 20.0, core_weighted=True, seed=7520),
    80: Workload("w80", 80, 1.5, 2500, 6.0, 20.0, laplace=True, core_weighted=True, seed=7720),
    128: Workload("w128", 128, 1.2, 6000, 5.0, 10.0, core_weighted=True, seed=3020),
    144: Workload("w144", 144, 1.2, 6000, 5.0, 20.0, core_weighted=True, seed=7820),
}


@pytest.mark.parametrize("n", sorted(WEIGHTED))
def test_fused_core_weighted_mask_vs_oracle(n):
    """Core-weighted masks take the three-grid forms of K1 and K2 (W^2 is its
    own grid) at the sizes whose other fused cases all use 0/1 masks."""
    w = WEIGHTED[n]
    c = make_case(w)
    assert not set(np.unique(c.mask)) <= {0.0, 1.0}
    R = rotations_for_angle(w.angle)[:24]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say(f"scan {n}^3 fused, core-weighted mask, 24 rotations")
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused", batch=8)
    assert corr.fused_available
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    err, dec = assert_scan_parity(corr.lcc, corr.rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


BINARY = {48: "f48", 80: "f80", 128: "search128", 144: "f144"}


@pytest.mark.parametrize("n", sorted(BINARY))
def test_fused_without_binary_shortcut_vs_oracle(n):
    """0/1 masks with the W^2 = W shortcut off: the same three-grid forms."""
    from test_gpu_scan import NEW_SIZES
    w = NEW_SIZES.get(n) or WORKLOADS[BINARY[n]]
    c = make_case(w)
    assert set(np.unique(c.mask)) <= {0.0, 1.0}
    R = rotations_for_angle(w.angle)[:16]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say(f"scan {n}^3 fused, binary shortcut off, 16 rotations")
    best, rot = run(c.target, c.template, c.mask, R, c.laplace, "fused", 8, binary_shortcut=False)
    err, dec = assert_scan_parity(best, rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


@pytest.mark.parametrize("mode,name", [("fused", "cpu64"), ("batched", "cpu64"),
                                       ("fused", "w48")])
def test_duplicate_rotations_keep_the_first_index(mode, name):
    """A rotation set listed twice, and each rotation repeated in place: the
    copies tie bit for bit with the originals, and strict '>' keeps the lower
    index everywhere (powerfitter.py:323-330), so the maps equal those of the
    single set exactly (indices mapped)."""
    w = WEIGHTED[48] if name == "w48" else WORKLOADS[name]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:12]
    say(f"scan {name} {mode}: 12 rotations, twice and repeated")
    one = run(c.target, c.template, c.mask, R, c.laplace, mode, 8)
    two = run(c.target, c.template, c.mask, np.concatenate([R, R]), c.laplace, mode, 8)
    rep = run(c.target, c.template, c.mask, np.repeat(R, 2, axis=0), c.laplace, mode, 5)
    assert np.array_equal(one[0], two[0]) and np.array_equal(one[1], two[1])
    assert np.array_equal(one[0], rep[0]) and np.array_equal(2 * one[1], rep[1])
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    assert_scan_parity(*one, ref)


@pytest.mark.parametrize("n,mode", [(32, "batched"), (64, "batched"), (64, "fused")])
def test_zero_variance_nan_never_wins(n, mode):
    """Constant target, +-1 template, cube-group rotations: G = 0 and the
    variance is exactly 0 at every voxel (the oracle divides 0 by 0 there),
    so no rotation may update the initial (LCC 0, index 0)."""
    target, template, mask = zero_variance_case(n)
    R, _ = cube_group()
    say(f"scan {n}^3 {mode}: zero-variance case, 24 cube rotations")
    ref = oc.scan(target, template, mask, R)
    assert not ref.best.any() and not ref.rot.any()
    best, rot = run(target, template, mask, R, False, mode, 8)
    assert not best.any() and not rot.any()


def test_graph_recaptured_after_inputs_change():
    """A captured scan, then a new rotation set of the same length, a new
    mask and a larger eager batch: every later replay equals the eager scan
    of the current inputs and the oracle."""
    c = make_case(WORKLOADS["cpu64"])
    R = rotations_for_angle(10.0)
    R1, R2 = R[:32], R[100:132]
    say("graph capture cpu64: replay after set_rotations / mask / larger batch")
    corr = HipCorrelator(c.target.astype(np.float64), mode="fused", batch=16)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R1
    corr.reset()
    corr.enqueue_scan_graph(0, 32)
    first = tuple(a.copy() for a in corr.result())
    corr.scan()
    assert np.array_equal(first[0], corr.lcc) and np.array_equal(first[1], corr.rot)

    def check(template, mask, Rs):
        corr.reset()
        corr.enqueue_scan_graph(0, 32)
        lcc, rot = (a.copy() for a in corr.result())
        corr.scan()
        assert np.array_equal(lcc, corr.lcc) and np.array_equal(rot, corr.rot)
        ref = oc.scan(c.target.astype(np.float64), template, mask, Rs)
        assert_scan_parity(lcc, rot, ref)

    corr.rotations = R2                  # same length, new array
    check(c.template, c.mask, R2)
    core = make_case(Workload("w64", 64, 2.0, 1500, 8.0, 10.0, core_weighted=True, seed=1020))
    corr.template = core.template        # new template and weighted mask
    corr.mask = core.mask
    check(core.template, core.mask, R2)
    corr.batch = 32                      # eager scan with a larger batch reallocates
    corr.scan()
    corr.batch = 16
    check(core.template, core.mask, R2)


def test_plain_path_config2_vs_oracle():
    """The 96^3 configuration with Laplace and core weighting on the plain
    one-rotation-per-launch path."""
    c = make_case(WORKLOADS["lc96"])
    R = rotations_for_angle(10.0)[:8]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say("scan lc96 plain, 8 rotations")
    best, rot = run(c.target, c.template, c.mask, R, c.laplace, "plain", 1)
    err, dec = assert_scan_parity(best, rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")
