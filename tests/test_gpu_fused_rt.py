"""GPU parity of the run-time planned fused pipeline (PF_MODE_FUSED_RT): grids
without a compiled fused size -- 120^3 (8 * 3 * 5) and the non-cubic,
radix-7 and odd-length shapes the powerfit command line produces (``extend``
to the next 2,3,5,7-smooth length per axis, powerfit.py:231-233) -- against
the FP64 oracle, with 0/1, core-weighted and Laplace-filtered inputs.

Bar: best LCC within 1e-4 of the oracle; rotation indices identical where
the oracle's winner leads by more than 1e-4.
"""
import os
import sys

import numpy as np
import pytest

from oracle import correlator as oc
from powerfit_amd import HipCorrelator
from powerfit_amd.rotations import rotations_for_angle
from powerfit_amd.synth import Workload, make_case

from parity import assert_scan_parity

pytestmark = pytest.mark.gpu
NPROC = min(8, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = {
    "c120": Workload("c120", 120, 1.5, 4000, 6.0, 20.0, seed=8100),
    "c120_lc": Workload("c120_lc", 120, 1.5, 4000, 6.0, 20.0, laplace=True, core_weighted=True,
                        seed=8101),
    "s84x96x112": Workload("s84x96x112", 84, 1.6, 3000, 6.5, 20.0, seed=8102,
                           shape=(84, 96, 112)),
    "s70x72x98_cw": Workload("s70x72x98_cw", 70, 1.8, 2000, 7.0, 20.0, core_weighted=True,
                             seed=8103, shape=(70, 72, 98)),
    "s50x60x42_l": Workload("s50x60x42_l", 42, 2.0, 800, 8.0, 20.0, laplace=True, seed=8104,
                            shape=(50, 60, 42)),
    # z = 44 and x / 2 = 33 carry a factor 11: no compiled transform, so the
    # z stage and the row passes take their run-time planned forms
    "s44x48x66": Workload("s44x48x66", 44, 2.0, 800, 8.0, 20.0, seed=8107,
                          shape=(44, 48, 66)),
    # odd y length (63 = 9 * 7) with Laplace and core weighting
    "s60x63x72_lc": Workload("s60x63x72_lc", 60, 2.0, 1200, 8.0, 20.0, laplace=True,
                             core_weighted=True, seed=8108, shape=(60, 63, 72)),
    # short z axis: two threads per z-column, blocks below a wavefront
    "s24x36x40": Workload("s24x36x40", 24, 2.5, 300, 9.0, 20.0, seed=8109, shape=(24, 36, 40)),
    # odd x lengths: rows y = 2p, 2p + 1 packed into one complex transform;
    # odd y as well (the last row unpaired), Laplace and core weighting
    "s60x64x75_cw": Workload("s60x64x75_cw", 60, 2.0, 1200, 8.0, 20.0, core_weighted=True,
                             seed=8110, shape=(60, 64, 75)),
    "s49x45x81_l": Workload("s49x45x81_l", 45, 2.0, 900, 8.0, 20.0, laplace=True, seed=8111,
                            shape=(49, 45, 81)),
    "s48x50x45": Workload("s48x50x45", 45, 2.0, 600, 8.0, 20.0, seed=8106, shape=(48, 50, 45)),
    # odd x length with a factor 11 (the run-time planned row pass)
    "s40x42x55": Workload("s40x42x55", 40, 2.0, 500, 8.0, 20.0, seed=8112, shape=(40, 42, 55)),
}


def scan(c, R, mode, batch=8, **kw):
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode=mode, batch=batch,
                         **kw)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    return corr


@pytest.mark.parametrize("name", sorted(CASES))
def test_fused_rt_vs_oracle(name):
    w = CASES[name]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:16]
    say(f"scan {name} {c.target.shape} fused_rt, 16 rotations")
    corr = scan(c, R, "fused_rt")
    assert corr.fused_rt_available and not corr.fused_available
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    err, dec = assert_scan_parity(corr.lcc, corr.rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


def test_auto_takes_fused_rt_and_batches_agree():
    """AUTO runs the run-time planned kernels where no compiled size exists;
    batch and chunk sizes and rotation ranges do not change a result."""
    c = make_case(CASES["s84x96x112"])
    R = rotations_for_angle(20.0)[:12]
    say("s84x96x112: auto vs fused_rt at batches 12 / 5, ranges [0, 7) + [7, 12)")
    a = scan(c, R, "auto", batch=12)
    b = scan(c, R, "fused_rt", batch=5)
    assert np.array_equal(a.lcc, b.lcc) and np.array_equal(a.rot, b.rot)
    b.set_tuning(1, 1)
    b.reset()
    b.enqueue_scan(0, 7)
    b.enqueue_scan(7, 5)
    lcc, rot = b.result()
    assert np.array_equal(lcc, a.lcc) and np.array_equal(rot, a.rot)


def test_fused_rt_without_binary_shortcut():
    """0/1 mask with W^2 transformed as its own grid (the weighted form)."""
    c = make_case(CASES["c120"])
    R = rotations_for_angle(20.0)[:8]
    say("c120 fused_rt, binary shortcut off, 8 rotations")
    corr = scan(c, R, "fused_rt", binary_shortcut=False)
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    assert_scan_parity(corr.lcc, corr.rot, ref)


@pytest.mark.parametrize("name", ["c120_lc", "s84x96x112", "s70x72x98_cw", "s60x64x75_cw",
                                  "s48x50x45"])
def test_alternative_z_register_plan_vs_oracle(name):
    """z schedule 1: the run-time planned z stage's alternative register plan
    (8 slots per thread) where the z length has one -- 120, 84, 70, 60 and 48
    here -- within FP32 rounding of the default plan and within the bar of
    the oracle."""
    w = CASES[name]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:8]
    say(f"scan {name} {c.target.shape} fused_rt, z schedule 1 vs 0, 8 rotations")
    a = scan(c, R, "fused_rt")
    b = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused_rt", batch=8)
    b.template = c.template
    b.mask = c.mask
    b.rotations = R
    b.set_schedule(1)
    b.scan()
    d = float(np.abs(a.lcc - b.lcc).max())
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    err, dec = assert_scan_parity(b.lcc, b.rot, ref)
    say(f"  |schedule 1 - schedule 0| {d:.2e}; max err {err:.2e}, decisive voxels {dec}")
    assert d <= 1e-5


def test_fused_rt_on_a_compiled_size_matches_oracle():
    """Forced onto a grid that also has compiled kernels (64^3): the same
    results within the tolerance."""
    c = make_case(Workload("rt64", 64, 2.0, 1500, 8.0, 10.0, seed=8105))
    R = rotations_for_angle(10.0)[:16]
    say("64^3 fused_rt vs compiled fused vs oracle, 16 rotations")
    rt = scan(c, R, "fused_rt")
    fu = scan(c, R, "fused")
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R, nproc=NPROC)
    assert_scan_parity(rt.lcc, rt.rot, ref)
    assert np.abs(rt.lcc - fu.lcc).max() <= 1e-4


def test_large_plane_falls_back_to_batched():
    """A plane of more than 16,384 voxels (130 x 130) is beyond the LCC stage's
    voxels per thread: AUTO runs the unfused kernels, fused_rt is refused."""
    c = make_case(Workload("big", 24, 2.0, 600, 8.0, 20.0, seed=8113, shape=(24, 130, 130)))
    R = rotations_for_angle(20.0)[:6]
    corr = scan(c, R, "auto")
    assert not corr.fused_rt_available and not corr.fused_available
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R, nproc=NPROC)
    assert_scan_parity(corr.lcc, corr.rot, ref)
    with pytest.raises(RuntimeError):
        scan(c, R, "fused_rt")


def test_odd_x_batches_and_ranges_agree():
    """Odd x: batch sizes, chunks and rotation ranges do not change a result,
    and AUTO takes the run-time planned pipeline."""
    c = make_case(CASES["s60x64x75_cw"])
    R = rotations_for_angle(20.0)[:10]
    a = scan(c, R, "auto", batch=10)
    assert a.fused_rt_available
    b = scan(c, R, "fused_rt", batch=3)
    assert np.array_equal(a.lcc, b.lcc) and np.array_equal(a.rot, b.rot)
    b.set_tuning(1, 1)
    b.reset()
    b.enqueue_scan(0, 4)
    b.enqueue_scan(4, 6)
    lcc, rot = b.result()
    assert np.array_equal(lcc, a.lcc) and np.array_equal(rot, a.rot)


def test_fused_rt_graph_and_noalloc_equal_eager():
    """The run-time planned stages under a captured graph and through the
    allocation-free entry point give the eager scan's maps bit for bit."""
    from powerfit_amd import _lib
    c = make_case(CASES["s84x96x112"])
    R = rotations_for_angle(20.0)[:10]
    say("s84x96x112 fused_rt: graph replays and pf_scan_noalloc vs eager")
    corr = scan(c, R, "fused_rt", batch=4)
    eager = (corr.lcc.copy(), corr.rot.copy())
    for _ in range(2):
        corr.reset()
        corr.enqueue_scan_graph(0, 10)
        lcc, rot = corr.result()
        assert np.array_equal(lcc, eager[0]) and np.array_equal(rot, eager[1])
    lib = _lib.load()
    corr.reset()
    _lib.check(lib.pf_scan_noalloc(corr._ctx, 0, 10, 4, _lib.MODES["fused_rt"]),
               "pf_scan_noalloc")
    lcc, rot = corr.result()
    assert np.array_equal(lcc, eager[0]) and np.array_equal(rot, eager[1])
