"""pf_scan_noalloc: the scan whose work buffers live in the library's static
per-GPU scratch instead of ctx allocations.  Called through the C-ABI and
checked against pf_scan (same kernels, same batch: bit-identical maps) and
against the FP64 oracle, on two calls in a row where the second has a
smaller input, with the unfused stages (real grids in the scratch too), and
with a batch too large for the scratch (rejected, nothing enqueued).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import correlator as oc
from powerfit_amd import HipCorrelator, _lib
from powerfit_amd.rotations import rotations_for_angle
from powerfit_amd.synth import WORKLOADS, Workload, make_case

from parity import assert_scan_parity

pytestmark = pytest.mark.gpu
NPROC = min(8, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def prepared(w, nrot):
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:nrot]
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    return c, R, corr


def scan_with(corr, entry, first, count, batch, mode):
    lib = _lib.load()
    corr.reset()
    _lib.check(getattr(lib, entry)(corr._ctx, first, count, batch, _lib.MODES[mode]), entry)
    return tuple(a.copy() for a in corr.result())


def check_pair(name, c, R, corr, batch, mode):
    say(f"{name}: pf_scan vs pf_scan_noalloc, {len(R)} rotations, batch {batch}, {mode}")
    ref_maps = scan_with(corr, "pf_scan", 0, len(R), batch, mode)
    got = scan_with(corr, "pf_scan_noalloc", 0, len(R), batch, mode)
    assert np.array_equal(got[0], ref_maps[0]) and np.array_equal(got[1], ref_maps[1])
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    err, dec = assert_scan_parity(*got, ref)
    say(f"  identical maps; max |dLCC| {err:.2e} against the oracle, decisive voxels {dec}")


def test_noalloc_equals_scan_then_smaller_input():
    # first call: 64^3 fused, 24 rotations in batches of 8
    c, R, corr = prepared(WORKLOADS["cpu64"], 24)
    check_pair("cpu64 fused", c, R, corr, 8, "fused")
    # second call, smaller input: 48^3 core-weighted (three-grid forms), 12 rotations
    w48 = Workload("w48na", 48, 2.0, 600, 8.0, 20.0, core_weighted=True, seed=7521)
    c2, R2, corr2 = prepared(w48, 12)
    check_pair("48^3 core-weighted fused", c2, R2, corr2, 4, "fused")


def test_noalloc_unfused_stages_and_ranges():
    # batched mode puts the real grids in the scratch after the spectra
    w = Workload("w40na", 40, 2.0, 500, 8.0, 20.0, laplace=True, seed=7522)
    c, R, corr = prepared(w, 10)
    check_pair("40^3 batched", c, R, corr, 4, "batched")
    # two rotation ranges merge like pf_scan's
    say("40^3 batched: ranges [0, 6) + [6, 10) through pf_scan_noalloc")
    lib = _lib.load()
    corr.reset()
    for first, count in ((0, 6), (6, 4)):
        _lib.check(lib.pf_scan_noalloc(corr._ctx, first, count, 3, _lib.MODES["batched"]),
                   "pf_scan_noalloc")
    split = tuple(a.copy() for a in corr.result())
    whole = scan_with(corr, "pf_scan", 0, 10, 4, "batched")
    assert np.array_equal(split[0], whole[0]) and np.array_equal(split[1], whole[1])


def test_noalloc_rejects_a_batch_beyond_the_scratch():
    n = 128
    rng = np.random.default_rng(7523)
    target = rng.random((n, n, n))
    mask = np.zeros((n, n, n))
    mask[:4, :4, :4] = 1.0
    template = rng.random((n, n, n)) * mask
    corr = HipCorrelator(target, mode="fused")
    corr.template = template
    corr.mask = mask
    corr.rotations = rotations_for_angle(20.0)[:8]
    lib = _lib.load()
    H = n * n * (n // 2 + 1)
    fits = (128 << 20) // (24 * H)         # rotations per launch the scratch holds
    assert fits == 5
    say(f"128^3 fused: batch {fits + 1} rejected, batch {fits} accepted")
    rc = lib.pf_scan_noalloc(corr._ctx, 0, 8, fits + 1, _lib.MODES["fused"])
    assert rc == -1
    assert b"scratch" in lib.pf_last_error()
    corr.reset()
    _lib.check(lib.pf_scan_noalloc(corr._ctx, 0, 8, fits, _lib.MODES["fused"]),
               "pf_scan_noalloc")
    got = tuple(a.copy() for a in corr.result())
    want = scan_with(corr, "pf_scan", 0, 8, fits, "fused")
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
